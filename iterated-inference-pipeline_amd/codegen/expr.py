"""Expression language of PLOM model files: tokenizer, parser and C emitter.

Rates, observation means/variances and white-noise intensities in process.json and
link.json are strings such as ``r0/N*v*(1.0+e*sin(2*M_PI*t/ONE_YEAR))*I``.  The
reference splits them on the operator set ``+ - * / , ( )``
(model_builder/Cmodel.py:33, ``change_user_input`` :188-209) and hands them to SymPy
(model_builder/Ccoder.py:104-139).  Here they are parsed into a small tree and printed
straight to C with an explicit, left-to-right evaluation order, so the same text
compiles to the same arithmetic under gcc (CPU oracle) and hipcc (gfx950).
"""
import re

OPS = set("+-*/,()")

_TOKEN_RE = re.compile(
    r"\s*(?:(?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?)|(?P<name>[A-Za-z_][A-Za-z_0-9]*)"
    r"|(?P<op>\*\*|\^|[-+*/(),]))")


def change_user_input(s):
    """Split an expression the way the reference does (Cmodel.py:188-209): whitespace
    removed, one token per operator character, maximal runs of other characters."""
    s = s.replace(" ", "")
    out, cur = [], ""
    for ch in s:
        if ch in OPS:
            if cur:
                out.append(cur)
                cur = ""
            out.append(ch)
        else:
            cur += ch
    if cur:
        out.append(cur)
    return out


def tokenize(s):
    pos, toks = 0, []
    s = s.strip()
    while pos < len(s):
        m = _TOKEN_RE.match(s, pos)
        if not m or m.end() == pos:
            raise ValueError("cannot tokenize %r at %d" % (s, pos))
        if m.group("num") is not None:
            toks.append(("num", m.group("num")))
        elif m.group("name") is not None:
            toks.append(("name", m.group("name")))
        else:
            op = m.group("op")
            toks.append(("op", "**" if op == "^" else op))
        pos = m.end()
    return toks


class Node:
    pass


class Num(Node):
    def __init__(self, text):
        self.text = text
        self.value = float(text)


class Sym(Node):
    def __init__(self, name):
        self.name = name


class Bin(Node):
    def __init__(self, op, a, b):
        self.op, self.a, self.b = op, a, b


class Neg(Node):
    def __init__(self, a):
        self.a = a


class Call(Node):
    def __init__(self, fn, args):
        self.fn, self.args = fn, args


class Parser:
    def __init__(self, text):
        self.text = text
        self.toks = tokenize(text)
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self, kind=None, val=None):
        t = self.peek()
        if t[0] is None or (kind and t[0] != kind) or (val and t[1] != val):
            raise ValueError("parse error in %r near token %d (%r)" % (self.text, self.i, t))
        self.i += 1
        return t

    def parse(self):
        e = self.expr()
        if self.i != len(self.toks):
            raise ValueError("trailing tokens in %r" % self.text)
        return e

    def expr(self):
        e = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            e = Bin(op, e, self.term())
        return e

    def term(self):
        e = self.unary()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            e = Bin(op, e, self.unary())
        return e

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return Neg(self.unary())
        if self.peek() == ("op", "+"):
            self.take()
            return self.unary()
        return self.power()

    def power(self):
        base = self.atom()
        if self.peek() == ("op", "**"):
            self.take()
            return Bin("**", base, self.unary())
        return base

    def atom(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            return Num(val)
        if kind == "name":
            self.take()
            if self.peek() == ("op", "("):
                self.take()
                args = []
                if self.peek() != ("op", ")"):
                    args.append(self.expr())
                    while self.peek() == ("op", ","):
                        self.take()
                        args.append(self.expr())
                self.take("op", ")")
                return Call(val, args)
            return Sym(val)
        if (kind, val) == ("op", "("):
            self.take()
            e = self.expr()
            self.take("op", ")")
            return e
        raise ValueError("unexpected token %r in %r" % (val, self.text))


def parse(text):
    return Parser(text).parse()


def symbols(node, acc=None):
    """Set of symbol names referenced by an expression tree (function names excluded)."""
    if acc is None:
        acc = set()
    if isinstance(node, Sym):
        acc.add(node.name)
    elif isinstance(node, Bin):
        symbols(node.a, acc)
        symbols(node.b, acc)
    elif isinstance(node, Neg):
        symbols(node.a, acc)
    elif isinstance(node, Call):
        for a in node.args:
            symbols(a, acc)
    return acc


def calls(node, acc=None):
    if acc is None:
        acc = []
    if isinstance(node, Bin):
        calls(node.a, acc)
        calls(node.b, acc)
    elif isinstance(node, Neg):
        calls(node.a, acc)
    elif isinstance(node, Call):
        acc.append(node)
        for a in node.args:
            calls(a, acc)
    return acc


def evaluate(node, env, funcs):
    """Evaluate a tree with Python floats (used by tests and the host simulator)."""
    if isinstance(node, Num):
        return node.value
    if isinstance(node, Sym):
        return env[node.name]
    if isinstance(node, Neg):
        return -evaluate(node.a, env, funcs)
    if isinstance(node, Bin):
        a, b = evaluate(node.a, env, funcs), evaluate(node.b, env, funcs)
        return {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b,
                "/": lambda: a / b, "**": lambda: a ** b}[node.op]()
    if isinstance(node, Call):
        return funcs[node.fn](*[evaluate(a, env, funcs) for a in node.args])
    raise TypeError(node)


class CEmitter:
    """Print a tree as a C expression for one precision.

    ``resolve(name)`` returns the C text of a symbol; ``call(fn, arg_texts, node)``
    returns the C text of a function call (or None to use the defaults below);
    ``div(node, a_text, b_text)`` the C text of a division a / b in double (or None for
    the plain quotient).  ``real`` is 'double' or 'float'.
    """

    def __init__(self, resolve, real="double", call=None, div=None):
        self.resolve = resolve
        self.real = real
        self.call_hook = call
        self.div_hook = div
        self.sfx = "" if real == "double" else "f"

    def lit(self, text):
        v = float(text)
        if self.real == "double":
            return repr(v) if ("." in repr(v) or "e" in repr(v)) else repr(v) + ".0"
        return ("%r" % v) + "f" if ("." in repr(v) or "e" in repr(v)) else repr(v) + ".0f"

    def emit(self, node):
        if isinstance(node, Num):
            return self.lit(node.text)
        if isinstance(node, Sym):
            return self.resolve(node.name)
        if isinstance(node, Neg):
            return "(-%s)" % self.emit(node.a)
        if isinstance(node, Bin):
            if node.op == "**":
                b = node.b
                if isinstance(b, Num) and b.value == int(b.value) and abs(b.value) <= 16:
                    return "plom_powi%s(%s, %d)" % (self.sfx, self.emit(node.a), int(b.value))
                return "plom_pow%s(%s, %s)" % (self.sfx, self.emit(node.a), self.emit(node.b))
            if node.op == "/" and self.real == "float":
                return "plom_fdivf(%s, %s)" % (self.emit(node.a), self.emit(node.b))
            if node.op == "/" and self.div_hook:
                r = self.div_hook(node, self.emit(node.a), self.emit(node.b))
                if r is not None:
                    return r
            return "(%s %s %s)" % (self.emit(node.a), node.op, self.emit(node.b))
        if isinstance(node, Call):
            args = [self.emit(a) for a in node.args]
            if self.call_hook:
                r = self.call_hook(node.fn, args, node)
                if r is not None:
                    return r
            fn = node.fn
            simple = {"sin": "plom_sin", "cos": "plom_cos", "exp": "plom_exp", "log": "plom_log",
                      "sqrt": "plom_sqrt", "pow": "plom_pow", "heaviside": "plom_heaviside",
                      "ramp": "plom_ramp", "slowstep": "plom_slowstep", "abs": "plom_fabs",
                      "fabs": "plom_fabs"}
            if fn in simple:
                return "%s%s(%s)" % (simple[fn], self.sfx, ", ".join(args))
            raise ValueError("unsupported function %r" % fn)
        raise TypeError(node)
