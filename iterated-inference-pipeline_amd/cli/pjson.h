/*
 * pjson.h -- minimal JSON document model for the `smc` command.
 *
 * The reference reads and writes its settings and theta through jansson
 * (model_builder/C/core/json.c); jansson is not part of this build, so this is a small
 * DOM of its own: objects keep insertion order, numbers keep whether they were written
 * as integers, and printing follows jansson's default compact style ("key": value,
 * separated by ", ") so a piped theta reads back the same.
 */
#ifndef PJSON_H
#define PJSON_H

#include <stdio.h>

typedef enum { PJ_NULL, PJ_TRUE, PJ_FALSE, PJ_INT, PJ_REAL, PJ_STRING, PJ_ARRAY, PJ_OBJECT } pj_type;

typedef struct pj {
    pj_type type;
    double num;          /* PJ_INT and PJ_REAL */
    char *str;           /* PJ_STRING */
    int n, cap;          /* PJ_ARRAY / PJ_OBJECT */
    struct pj **items;
    char **keys;         /* PJ_OBJECT */
} pj;

pj *pj_parse(const char *text, char *err, int errlen);
pj *pj_load_file(const char *path, char *err, int errlen);
void pj_free(pj *v);

pj *pj_get(const pj *obj, const char *key);       /* NULL if absent or not an object */
pj *pj_at(const pj *arr, int i);                  /* NULL if out of range */
int pj_size(const pj *v);
double pj_number(const pj *v, int *ok);
const char *pj_string(const pj *v);
int pj_is_number(const pj *v);

pj *pj_new_real(double x);
pj *pj_new_string(const char *s);
pj *pj_new_object(void);
void pj_set(pj *obj, const char *key, pj *val);   /* replaces (and frees) an existing value */

void pj_dump(const pj *v, FILE *f);

#endif
