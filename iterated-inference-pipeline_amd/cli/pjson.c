/* pjson.c -- minimal JSON parser / printer (see pjson.h). */
#include "pjson.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const char *s;
    int pos;
    char *err;
    int errlen;
} pj_in;

static void fail(pj_in *in, const char *msg)
{
    if (in->err && in->err[0] == 0) snprintf(in->err, (size_t)in->errlen, "json: %s at offset %d", msg, in->pos);
}

static pj *new_node(pj_type t)
{
    pj *v = calloc(1, sizeof(pj));
    if (v) v->type = t;
    return v;
}

static void push(pj *c, const char *key, pj *v)
{
    if (c->n == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 8;
        c->items = realloc(c->items, (size_t)c->cap * sizeof(pj *));
        if (c->type == PJ_OBJECT) c->keys = realloc(c->keys, (size_t)c->cap * sizeof(char *));
    }
    c->items[c->n] = v;
    if (c->type == PJ_OBJECT) c->keys[c->n] = strdup(key);
    c->n++;
}

static void ws(pj_in *in)
{
    while (in->s[in->pos] && isspace((unsigned char)in->s[in->pos])) in->pos++;
}

static pj *parse_value(pj_in *in, int depth);

static char *parse_str(pj_in *in)
{
    if (in->s[in->pos] != '"') { fail(in, "expected string"); return NULL; }
    in->pos++;
    size_t cap = 32, n = 0;
    char *out = malloc(cap);
    while (in->s[in->pos] && in->s[in->pos] != '"') {
        char c = in->s[in->pos++];
        if (c == '\\') {
            char e = in->s[in->pos++];
            switch (e) {
            case 'n': c = '\n'; break;
            case 't': c = '\t'; break;
            case 'r': c = '\r'; break;
            case 'b': c = '\b'; break;
            case 'f': c = '\f'; break;
            case 'u': {
                unsigned cp = 0;
                for (int k = 0; k < 4; k++) {
                    char h = in->s[in->pos++];
                    cp = cp * 16 + (unsigned)(isdigit((unsigned char)h) ? h - '0' : (tolower((unsigned char)h) - 'a' + 10));
                }
                if (cp < 0x80) c = (char)cp;
                else { /* keep non-ASCII as UTF-8 */
                    char buf[4];
                    int m = 0;
                    if (cp < 0x800) { buf[m++] = (char)(0xC0 | (cp >> 6)); buf[m++] = (char)(0x80 | (cp & 0x3F)); }
                    else { buf[m++] = (char)(0xE0 | (cp >> 12)); buf[m++] = (char)(0x80 | ((cp >> 6) & 0x3F)); buf[m++] = (char)(0x80 | (cp & 0x3F)); }
                    for (int k = 0; k < m; k++) {
                        if (n + 2 >= cap) out = realloc(out, cap *= 2);
                        out[n++] = buf[k];
                    }
                    continue;
                }
                break;
            }
            default: c = e; break;
            }
        }
        if (n + 2 >= cap) out = realloc(out, cap *= 2);
        out[n++] = c;
    }
    if (in->s[in->pos] != '"') { fail(in, "unterminated string"); free(out); return NULL; }
    in->pos++;
    out[n] = 0;
    return out;
}

static pj *parse_value(pj_in *in, int depth)
{
    if (depth > 512) { fail(in, "nesting too deep"); return NULL; }
    ws(in);
    char c = in->s[in->pos];
    if (c == '{') {
        in->pos++;
        pj *o = new_node(PJ_OBJECT);
        ws(in);
        if (in->s[in->pos] == '}') { in->pos++; return o; }
        for (;;) {
            ws(in);
            char *k = parse_str(in);
            if (!k) { pj_free(o); return NULL; }
            ws(in);
            if (in->s[in->pos] != ':') { fail(in, "expected ':'"); free(k); pj_free(o); return NULL; }
            in->pos++;
            pj *v = parse_value(in, depth + 1);
            if (!v) { free(k); pj_free(o); return NULL; }
            push(o, k, v);
            free(k);
            ws(in);
            if (in->s[in->pos] == ',') { in->pos++; continue; }
            if (in->s[in->pos] == '}') { in->pos++; return o; }
            fail(in, "expected ',' or '}'");
            pj_free(o);
            return NULL;
        }
    }
    if (c == '[') {
        in->pos++;
        pj *a = new_node(PJ_ARRAY);
        ws(in);
        if (in->s[in->pos] == ']') { in->pos++; return a; }
        for (;;) {
            pj *v = parse_value(in, depth + 1);
            if (!v) { pj_free(a); return NULL; }
            push(a, NULL, v);
            ws(in);
            if (in->s[in->pos] == ',') { in->pos++; continue; }
            if (in->s[in->pos] == ']') { in->pos++; return a; }
            fail(in, "expected ',' or ']'");
            pj_free(a);
            return NULL;
        }
    }
    if (c == '"') {
        char *s = parse_str(in);
        if (!s) return NULL;
        pj *v = new_node(PJ_STRING);
        v->str = s;
        return v;
    }
    if (!strncmp(in->s + in->pos, "true", 4)) { in->pos += 4; return new_node(PJ_TRUE); }
    if (!strncmp(in->s + in->pos, "false", 5)) { in->pos += 5; return new_node(PJ_FALSE); }
    if (!strncmp(in->s + in->pos, "null", 4)) { in->pos += 4; return new_node(PJ_NULL); }
    if (c == '-' || isdigit((unsigned char)c)) {
        char *end = NULL;
        const char *start = in->s + in->pos;
        double x = strtod(start, &end);
        if (end == start) { fail(in, "bad number"); return NULL; }
        int is_int = 1;
        for (const char *p = start; p < end; p++)
            if (*p == '.' || *p == 'e' || *p == 'E') is_int = 0;
        pj *v = new_node(is_int ? PJ_INT : PJ_REAL);
        v->num = x;
        in->pos += (int)(end - start);
        return v;
    }
    fail(in, "unexpected character");
    return NULL;
}

pj *pj_parse(const char *text, char *err, int errlen)
{
    pj_in in = {text, 0, err, errlen};
    if (err && errlen) err[0] = 0;
    pj *v = parse_value(&in, 0);
    if (v) {
        ws(&in);
        if (in.s[in.pos]) { fail(&in, "trailing characters"); pj_free(v); return NULL; }
    }
    return v;
}

pj *pj_load_file(const char *path, char *err, int errlen)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        if (err) snprintf(err, (size_t)errlen, "cannot open %s", path);
        return NULL;
    }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = malloc((size_t)n + 1);
    size_t got = fread(buf, 1, (size_t)n, f);
    fclose(f);
    buf[got] = 0;
    pj *v = pj_parse(buf, err, errlen);
    free(buf);
    return v;
}

void pj_free(pj *v)
{
    if (!v) return;
    for (int i = 0; i < v->n; i++) {
        pj_free(v->items[i]);
        if (v->keys) free(v->keys[i]);
    }
    free(v->items);
    free(v->keys);
    free(v->str);
    free(v);
}

pj *pj_get(const pj *o, const char *key)
{
    if (!o || o->type != PJ_OBJECT) return NULL;
    for (int i = 0; i < o->n; i++)
        if (!strcmp(o->keys[i], key)) return o->items[i];
    return NULL;
}

pj *pj_at(const pj *a, int i)
{
    if (!a || (a->type != PJ_ARRAY && a->type != PJ_OBJECT) || i < 0 || i >= a->n) return NULL;
    return a->items[i];
}

int pj_size(const pj *v) { return (v && (v->type == PJ_ARRAY || v->type == PJ_OBJECT)) ? v->n : 0; }

int pj_is_number(const pj *v) { return v && (v->type == PJ_INT || v->type == PJ_REAL); }

double pj_number(const pj *v, int *ok)
{
    if (pj_is_number(v)) { if (ok) *ok = 1; return v->num; }
    if (ok) *ok = 0;
    return NAN;
}

const char *pj_string(const pj *v) { return (v && v->type == PJ_STRING) ? v->str : NULL; }

pj *pj_new_real(double x)
{
    pj *v = new_node(PJ_REAL);
    v->num = x;
    return v;
}

pj *pj_new_string(const char *s)
{
    pj *v = new_node(PJ_STRING);
    v->str = strdup(s);
    return v;
}

pj *pj_new_object(void) { return new_node(PJ_OBJECT); }

void pj_set(pj *o, const char *key, pj *val)
{
    for (int i = 0; i < o->n; i++)
        if (!strcmp(o->keys[i], key)) {
            pj_free(o->items[i]);
            o->items[i] = val;
            return;
        }
    push(o, key, val);
}

static void dump_str(const char *s, FILE *f)
{
    fputc('"', f);
    for (; *s; s++) {
        unsigned char c = (unsigned char)*s;
        if (c == '"' || c == '\\') fprintf(f, "\\%c", c);
        else if (c == '\n') fputs("\\n", f);
        else if (c == '\t') fputs("\\t", f);
        else if (c < 0x20) fprintf(f, "\\u%04x", c);
        else fputc(c, f);
    }
    fputc('"', f);
}

void pj_dump(const pj *v, FILE *f)
{
    char buf[64];
    switch (v->type) {
    case PJ_NULL: fputs("null", f); break;
    case PJ_TRUE: fputs("true", f); break;
    case PJ_FALSE: fputs("false", f); break;
    case PJ_INT: fprintf(f, "%.0f", v->num); break;
    case PJ_REAL:
        snprintf(buf, sizeof buf, "%.17g", v->num);
        if (!strpbrk(buf, ".eEn")) strcat(buf, ".0");
        fputs(buf, f);
        break;
    case PJ_STRING: dump_str(v->str, f); break;
    case PJ_ARRAY:
        fputc('[', f);
        for (int i = 0; i < v->n; i++) {
            if (i) fputs(", ", f);
            pj_dump(v->items[i], f);
        }
        fputc(']', f);
        break;
    case PJ_OBJECT:
        fputc('{', f);
        for (int i = 0; i < v->n; i++) {
            if (i) fputs(", ", f);
            dump_str(v->keys[i], f);
            fputs(": ", f);
            pj_dump(v->items[i], f);
        }
        fputc('}', f);
        break;
    }
}
