"""Writes tests/golden/presets/<id>.json: the reference's new-file presets as data.

The presets (src/editor/presets/preset-*.ts in the reference) are TypeScript object literals built
with the helpers of presets/types.ts (createScalarEmitter / createVelocityEmitter /
createBurstEmitter / createIgniterEmitter add a `kind` to the props; createLight pairs a label
with its props) and, in one file, a local helper function returning an array spread into
`emitters`. This script evaluates that literal subset (objects, arrays, strings and template
strings, numbers, + and -, calls of those helpers and of functions declared in the same file,
spreads, shorthand properties) and writes the resulting NewFilePreset objects (types.ts:16-25)
as JSON. It reads the reference tree given on the command line (default /root/reference) and is
run once, in the container that holds it; the JSON files are the fixtures.

Usage: python tests/golden/make_preset_fixtures.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "presets")

KIND_HELPERS = {"createScalarEmitter": "scalar", "createVelocityEmitter": "velocity",
                "createBurstEmitter": "burst", "createIgniterEmitter": "igniter"}

TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*|/\*.*?\*/)
  | (?P<num>\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)
  | (?P<str>'(?:[^'\\]|\\.)*'|"(?:[^"\\]|\\.)*")
  | (?P<tpl>`(?:[^`\\]|\\.)*`)
  | (?P<ident>[A-Za-z_$][A-Za-z0-9_$]*)
  | (?P<punct>\.\.\.|=>|[{}\[\](),:;=+\-*/<>|&?.])
""", re.S | re.X)


def tokenize(text):
    out, pos = [], 0
    while pos < len(text):
        m = TOKEN.match(text, pos)
        if not m:
            raise SyntaxError(f"unexpected {text[pos:pos + 30]!r}")
        pos = m.end()
        if m.lastgroup != "ws":
            out.append((m.lastgroup, m.group()))
    return out


class Parser:
    def __init__(self, toks, functions, env=None):
        self.t, self.i, self.functions, self.env = toks, 0, functions, dict(env or {})

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.t[self.i]
        if val is not None and tok[1] != val:
            raise SyntaxError(f"expected {val!r}, got {tok[1]!r}")
        self.i += 1
        return tok

    def expr(self):
        v = self.unary()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            r = self.unary()
            v = v + r if op == "+" else v - r
        return v

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            return -self.unary()
        return self.primary()

    def primary(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            return float(val) if any(c in val for c in ".eE") else int(val)
        if kind == "str":
            self.take()
            return bytes(val[1:-1], "utf-8").decode("unicode_escape")
        if kind == "tpl":
            self.take()
            return re.sub(r"\$\{([A-Za-z_$][A-Za-z0-9_$]*)\}", lambda m: str(self.env[m.group(1)]), val[1:-1])
        if val == "{":
            return self.obj()
        if val == "[":
            return self.arr()
        if val == "(":
            self.take("(")
            v = self.expr()
            self.take(")")
            return v
        if kind == "ident":
            self.take()
            if val in ("true", "false"):
                return val == "true"
            if self.peek()[1] == "(":
                return self.call(val)
            return self.env[val]
        raise SyntaxError(f"unexpected token {val!r}")

    def args(self):
        self.take("(")
        out = []
        while self.peek()[1] != ")":
            out.append(self.expr())
            if self.peek()[1] == ",":
                self.take()
        self.take(")")
        return out

    def call(self, name):
        a = self.args()
        if name in KIND_HELPERS:
            return {"label": a[0], "props": {"kind": KIND_HELPERS[name], **a[1]}}
        if name == "createLight":
            return {"label": a[0], "props": a[1]}
        params, body = self.functions[name]
        return Parser(body, self.functions, dict(zip(params, a))).expr()

    def obj(self):
        self.take("{")
        out = {}
        while self.peek()[1] != "}":
            if self.peek()[1] == "...":
                self.take()
                out.update(self.expr())
            else:
                kind, key = self.take()
                key = key[1:-1] if kind == "str" else key
                if self.peek()[1] == ":":
                    self.take()
                    out[key] = self.expr()
                else:  # shorthand property
                    out[key] = self.env[key]
            if self.peek()[1] == ",":
                self.take()
        self.take("}")
        return out

    def arr(self):
        self.take("[")
        out = []
        while self.peek()[1] != "]":
            if self.peek()[1] == "...":
                self.take()
                out.extend(self.expr())
            else:
                out.append(self.expr())
            if self.peek()[1] == ",":
                self.take()
        self.take("]")
        return out


def local_functions(toks):
    """`function name(p: T, ...): R { return <expr> }` -> {name: (params, tokens of <expr>)}."""
    fns, i = {}, 0
    while i < len(toks):
        if toks[i] == ("ident", "function"):
            name = toks[i + 1][1]
            j = i + 3
            params = []
            while toks[j][1] != ")":
                if toks[j][0] == "ident" and toks[j + 1][1] == ":":
                    params.append(toks[j][1])
                j += 1
            while toks[j][1] != "{":
                j += 1
            assert toks[j + 1] == ("ident", "return"), name
            k, depth = j + 2, 0
            start = k
            while True:  # the returned expression ends at the body's closing brace
                v = toks[k][1]
                if v in "([{":
                    depth += 1
                elif v in ")]}":
                    if depth == 0:
                        break
                    depth -= 1
                k += 1
            fns[name] = (params, toks[start:k])
            i = k
        i += 1
    return fns


def preset_literal(toks):
    """Tokens of the object literal after `export const <x>: NewFilePreset =`."""
    for i in range(len(toks) - 3):
        if toks[i] == ("ident", "NewFilePreset") and toks[i + 1][1] == "=":
            return toks[i + 2:]
    raise SyntaxError("no NewFilePreset literal")


def main(ref):
    src = os.path.join(ref, "src", "editor", "presets")
    os.makedirs(OUT, exist_ok=True)
    order = re.findall(r"^\s+(\w+Preset),$", open(os.path.join(src, "newFilePresets.ts")).read(), re.M)
    ids = []
    for fname in sorted(os.listdir(src)):
        if not fname.startswith("preset-"):
            continue
        toks = tokenize(open(os.path.join(src, fname)).read())
        preset = Parser(preset_literal(toks), local_functions(toks)).expr()
        preset["source"] = f"src/editor/presets/{fname}"
        with open(os.path.join(OUT, preset["id"] + ".json"), "w") as fh:
            json.dump(preset, fh, indent=1)
            fh.write("\n")
        ids.append((fname, preset["id"]))
    # the editor's menu order (newFilePresets.ts:20-35)
    names = {f: i for f, i in ids}
    menu = []
    for var in order:
        fname = "preset-" + re.sub(r"(?<!^)(?=[A-Z])", "-", var[:-len("Preset")]).lower() + ".ts"
        menu.append(names[fname])
    with open(os.path.join(OUT, "index.json"), "w") as fh:
        json.dump({"presets": menu, "source": "src/editor/presets/newFilePresets.ts"}, fh, indent=1)
        fh.write("\n")
    print(f"wrote {len(ids)} presets to {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
