"""Source converter: an MPI C program -> device code for the gpumpi runtime.

The reference converter (source_converter/converter.cpp:59-442, a clang libTooling tool)
rewrites a C MPI program so that nvcc compiles it for the reference's device MPI
library.  This one produces the same output contract for hipcc and the gpumpi runtime,
from clang's JSON AST dump (`clang -Xclang -ast-dump=json`; the ROCm clang is enough,
no libTooling headers are needed) plus text edits at the AST's source offsets:

  * every function written in the project gets `__device__ ` (converter.cpp:77-81);
  * `main` becomes `__gpu_main(int argc, char** argv)` and is followed by
    `__device__ int __gpu_max_threads() { return N; }` (converter.cpp:64-75);
  * old-style (K&R) parameter lists become prototypes (converter.cpp:84-115);
  * implicit pointer bit-casts, legal in C and not in C++, are made explicit
    (converter.cpp:117-121);
  * every variable with global storage (file scope or `static` local) becomes a
    per-rank array `__device__ T name[N]` with its C initializer repeated N times, and
    every use becomes `__gpu_global(name)`, which indexes the array by rank id
    (converter.cpp:122-177).  Const-qualified globals are left shared (one copy, no
    index): every rank sees the same value, and they stay usable in constant
    expressions, where the reference's indexed form does not compile;
  * identifiers named `class` get a `__decl_` prefix (converter.cpp:178-183);
  * with --malloc-version V > 0: the coalescing rewrite of malloc / MPI_Alltoall calls
    that are not under a branch or loop -- `__coalesced` threaded through a
    `bool __coalesced = false` parameter of each function that makes such a call,
    `true` in main, MPI_Alltoall renamed MPI_Alltoall_coalesce; for V = 3 and 7 main
    also brackets its body with __gpu_init_malloc() / __gpu_clean_malloc()
    (converter.cpp:184-239, 275-330) -- and `#define GPUMPI_MALLOC_V<V>` on top.

Only files inside the project directory are rewritten (converter.cpp:375-396); each
gets `#include <gpumpi/global_vars.h>` as its first line (converter.cpp:364).  N is
GPU_MPI_MAX_RANKS (default 1024, converter.cpp:42-57).

While parsing, `#include "mpi.h"` resolves to include/gpumpi/c/mpi.h, a C rendering of
the gpumpi API.
"""
import json
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
C_MPI_DIR = os.path.join(ROOT, "include", "gpumpi", "c")


def clang_path():
    for c in (os.environ.get("GPUMPI_CLANG"), "/opt/rocm/llvm/bin/clang", shutil.which("clang")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("clang not found: the converter parses C through clang's JSON AST")


def max_ranks():
    v = os.environ.get("GPU_MPI_MAX_RANKS")
    if v is None:
        return 1024
    n = int(v)
    if n <= 0:
        raise ValueError("GPU_MPI_MAX_RANKS must hold a positive number of ranks")
    return n


def parse(path, clang_args=()):
    cmd = [clang_path(), "-x", "c", "-fsyntax-only", "-Xclang", "-ast-dump=json",
           "-I", C_MPI_DIR, *clang_args, path]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"clang could not parse {path}:\n{r.stderr}")
    return json.loads(r.stdout)


# ----------------------------------------------------------------------------- locations
class Locator:
    """Resolves the JSON dumper's locations to (file, offset).  The dumper prints a
    location's file only when it differs from the previously printed one, so the walk
    must visit locations in dump order (JSON key order) to know each one's file."""

    def __init__(self, main_file):
        self.cur = main_file

    def bare(self, loc):
        if not loc or "offset" not in loc:
            return None
        if "file" in loc:
            self.cur = loc["file"]
        return (self.cur, loc["offset"], loc.get("tokLen", 0))

    def visit(self, loc):
        """Returns {'spelling': (f, off, len), 'expansion': ..., 'macro_arg': bool}."""
        if not loc:
            return None
        if "spellingLoc" in loc or "expansionLoc" in loc:
            sp = self.bare(loc.get("spellingLoc"))
            ex = self.bare(loc.get("expansionLoc"))
            arg = loc.get("isMacroArgExpansion") or (loc.get("expansionLoc") or {}).get("isMacroArgExpansion")
            return {"spelling": sp, "expansion": ex, "macro_arg": bool(arg)}
        b = self.bare(loc)
        return {"spelling": b, "expansion": b, "macro_arg": False}


def file_loc(l):
    """Where an insertion for this location goes: a macro argument's own text, else the
    expansion point (converter.cpp:79 getFileLoc)."""
    if l is None:
        return None
    return l["spelling"] if l["macro_arg"] else l["expansion"]


# ----------------------------------------------------------------------------- edits
class Edits:
    """Insertions and replacements at byte offsets of the original files, applied in one
    forward pass.  At one offset, insertions go in rank order (a cast before the
    __gpu_global( that it casts), and before a replacement that starts there."""

    def __init__(self):
        self.ops = {}    # file -> list of (start, end, kind, rank, seq, text)
        self.seen = set()
        self.seq = 0

    def insert(self, pos, text, rank=5, key=None):
        if pos is None:
            return
        k = (pos[0], pos[1], text, key)
        if k in self.seen:
            return
        self.seen.add(k)
        self.seq += 1
        self.ops.setdefault(pos[0], []).append((pos[1], pos[1], 0, rank, self.seq, text))

    def replace(self, f, start, end, text):
        k = (f, start, end, "R")
        if k in self.seen:
            return
        self.seen.add(k)
        self.seq += 1
        self.ops.setdefault(f, []).append((start, end, 1, 0, self.seq, text))

    def apply(self, f, data):
        out, pos = [], 0
        for (s, e, kind, rank, seq, text) in sorted(self.ops.get(f, []), key=lambda o: o[:5]):
            if s < pos:
                continue  # inside a replaced span
            out.append(data[pos:s])
            out.append(text.encode())
            pos = e if kind == 1 else s
        out.append(data[pos:])
        return b"".join(out)


# ----------------------------------------------------------------------------- converter
LOOP_OR_BRANCH = {"IfStmt", "ForStmt", "WhileStmt", "DoStmt", "ConditionalOperator",
                  "BinaryConditionalOperator", "SwitchStmt"}


class Converter:
    def __init__(self, main_file, project_dir, malloc_version=0, ranks=None):
        self.main_file = main_file
        self.project_dir = os.path.realpath(project_dir)
        self.malloc_version = malloc_version
        self.N = ranks if ranks is not None else max_ranks()
        self.edits = Edits()
        self.files = {}  # path as dumped -> real path, for project files
        self.global_ids = set()      # VarDecl ids to index by rank
        self.const_global_ids = set()
        self.class_ids = set()
        self.funcs = {}              # name -> FunctionDecl info (project definitions)

    def in_project(self, f):
        if f is None:
            return False
        if f not in self.files:
            rp = os.path.realpath(f)
            self.files[f] = rp if (rp == os.path.realpath(self.main_file) or
                                   rp.startswith(self.project_dir + os.sep)) else None
        return self.files[f] is not None

    def text(self, f):
        return open(self.files.get(f) or f, "rb").read()

    # -- pass 1: walk in dump order, resolving every location
    def resolve(self, node, loc, parent_kind):
        if isinstance(node, dict):
            node["_parent"] = parent_kind
            for k, v in list(node.items()):
                if k == "loc":
                    node["_loc"] = loc.visit(v)
                elif k == "range":
                    node["_begin"] = loc.visit(v.get("begin"))
                    node["_end"] = loc.visit(v.get("end"))
                elif k == "inner":
                    for c in v:
                        self.resolve(c, loc, node.get("kind"))
                elif isinstance(v, dict) and any(x in v for x in ("offset", "spellingLoc", "expansionLoc")):
                    loc.visit(v)
        return node

    def walk(self, node, fn):
        stack = [node]
        while stack:
            n = stack.pop()
            if isinstance(n, dict):
                fn(n)
                for c in reversed(n.get("inner", [])):
                    stack.append(c)

    def project_node(self, n):
        b = file_loc(n.get("_begin")) or file_loc(n.get("_loc"))
        return b is not None and self.in_project(b[0])

    # -- the rules
    def convert(self, ast):
        sys.setrecursionlimit(max(10000, sys.getrecursionlimit()))
        self.resolve(ast, Locator(self.main_file), None)
        self.walk(ast, self.collect)
        self.walk(ast, self.rewrite)
        if self.malloc_version > 0:
            self.coalesce(ast)
        out = {}
        for f, real in self.files.items():
            if real is None:
                continue
            header = b"#include <gpumpi/global_vars.h>\n"
            if self.malloc_version > 0:
                header = f"#define GPUMPI_MALLOC_V{self.malloc_version}\n".encode() + header
            out[real] = header + self.edits.apply(f, self.text(f))
        if os.path.realpath(self.main_file) not in out:
            out[os.path.realpath(self.main_file)] = b"#include <gpumpi/global_vars.h>\n" + open(self.main_file, "rb").read()
        return out

    def collect(self, n):
        k = n.get("kind")
        if k == "VarDecl" and self.project_node(n):
            top = n["_parent"] == "TranslationUnitDecl"
            static_local = (not top) and n.get("storageClass") == "static"
            if top or static_local:
                if self.is_const_object(n.get("type", {}).get("qualType", "")):
                    self.const_global_ids.add(n["id"])
                else:
                    self.global_ids.add(n["id"])
        if n.get("name") == "class" and k in ("VarDecl", "ParmVarDecl", "FieldDecl", "FunctionDecl") \
                and self.project_node(n):
            self.class_ids.add(n["id"])
        if k == "FunctionDecl" and self.project_node(n) and not n.get("isImplicit"):
            self.funcs.setdefault(n["name"], []).append(n)

    @staticmethod
    def is_const_object(qual_type):
        """Whether the variable itself (or every array element) is const: `const int[4]`,
        `char *const`, not `const char *[3]` (mutable pointers to const chars)."""
        t = re.sub(r"(\s*\[\d*\])+$", "", qual_type).strip()
        if "*" in t:
            return t.endswith("const")
        return t.startswith("const ") or " const" in t

    def name_end(self, n):
        l = n.get("_loc")
        p = file_loc(l)
        return None if p is None else (p[0], p[1] + p[2])

    def rewrite(self, n):
        k = n.get("kind")
        if k == "FunctionDecl" and not n.get("isImplicit") and self.project_node(n):
            self.rewrite_function(n)
        elif k == "VarDecl" and (n["id"] in self.global_ids or n["id"] in self.const_global_ids):
            self.rewrite_global(n)
        elif k == "DeclRefExpr":
            ref = n.get("referencedDecl", {})
            if ref.get("id") in self.global_ids:
                b = n["_begin"]["spelling"] if n.get("_begin") else None
                e = n["_end"]["spelling"] if n.get("_end") else None
                if b and e and self.in_project(b[0]):
                    self.edits.insert(b, "__gpu_global(", rank=6, key="gg")
                    self.edits.insert((e[0], e[1] + e[2]), ")", rank=4, key="gg")
            if ref.get("id") in self.class_ids or (ref.get("name") == "class" and ref.get("kind") != "FunctionDecl"
                                                   and self.project_node(n)):
                p = file_loc(n.get("_begin"))
                if p and self.in_project(p[0]):
                    self.edits.insert(p, "__decl_", rank=7, key="cls")
        elif k == "MemberExpr" and n.get("name") == "class" and self.project_node(n):
            l = file_loc(n.get("_loc"))
            if l:
                self.edits.insert(l, "__decl_", rank=7, key="cls")
        elif k == "ImplicitCastExpr" and n.get("castKind") == "BitCast" and self.project_node(n):
            p = file_loc(n.get("_begin"))
            if p and self.in_project(p[0]):
                self.edits.insert(p, "(" + n["type"]["qualType"] + ")", rank=1, key="cast")
        if n.get("id") in self.class_ids and k != "DeclRefExpr":
            l = file_loc(n.get("_loc"))
            if l:
                self.edits.insert(l, "__decl_", rank=7, key="cls")

    def parens(self, f, name_end_off):
        """Offsets of '(' and its matching ')' after a function name."""
        data = self.text(f)
        i = name_end_off
        while i < len(data) and data[i:i + 1] != b"(":
            i += 1
        depth, j = 0, i
        while j < len(data):
            c = data[j:j + 1]
            if c == b"(":
                depth += 1
            elif c == b")":
                depth -= 1
                if depth == 0:
                    return i, j
            j += 1
        return i, None

    def rewrite_function(self, n):
        begin = file_loc(n["_begin"])
        if begin is None:
            return
        self.edits.insert(begin, "__device__ ", rank=0, key="dev")
        name_loc = file_loc(n.get("_loc"))
        f = name_loc[0]
        lp, rp = self.parens(f, name_loc[1] + name_loc[2])
        if n.get("name") == "main":
            self.edits.replace(f, name_loc[1], name_loc[1] + name_loc[2], "__gpu_main")
            if rp is not None:
                self.edits.replace(f, lp, rp + 1, "(int argc, char** argv)")
            end = file_loc(n["_end"])
            if end is not None and any(c.get("kind") == "CompoundStmt" for c in n.get("inner", [])):
                self.edits.insert((end[0], end[1] + end[2]),
                                  "\n__device__ int __gpu_max_threads() {\n    return %d;\n}\n" % self.N,
                                  rank=9, key="maxthreads")
            return
        # K&R parameter declarations after the parameter list
        params = [c for c in n.get("inner", []) if c.get("kind") == "ParmVarDecl"]
        if rp is None or not params:
            return
        outside = [p for p in params if file_loc(p.get("_loc")) and file_loc(p["_loc"])[1] > rp]
        if not outside:
            return
        body = next((c for c in n.get("inner", []) if c.get("kind") == "CompoundStmt"), None)
        proto = ", ".join(f'{p["type"]["qualType"]} {p.get("name", "")}'.strip() for p in params)
        self.edits.replace(f, lp, rp + 1, "(" + proto + ")")
        if body is not None:
            bstart = file_loc(body["_begin"])[1]
            self.edits.replace(f, rp + 1, bstart, " ")

    def rewrite_global(self, n):
        begin = file_loc(n["_begin"])
        if begin is None:
            return
        self.edits.insert(begin, "__device__ ", rank=0, key="dev")
        if n["id"] in self.const_global_ids:
            return
        ne = self.name_end(n)
        self.edits.insert(ne, "[%d]" % self.N, rank=2, key="dim")
        # `T name[] = {...}`: only the outer dimension of the per-rank array may be left
        # open, so the size the initializer gave goes into the brackets
        m = re.match(rb"\s*\[\s*\]", self.text(ne[0])[ne[1]:ne[1] + 64])
        size = re.search(r"\[(\d+)\]", n.get("type", {}).get("qualType", ""))
        if m and size:
            self.edits.replace(ne[0], ne[1], ne[1] + m.end(), "[%s]" % size.group(1))
        if n.get("init") and n.get("inner"):
            if n["init"] != "c":
                raise RuntimeError(f"only C-style initialization is supported for global {n.get('name')}")
            init = n["inner"][-1]
            b, e = file_loc(init["_begin"]), file_loc(init["_end"])
            if b and e and b[0] == e[0]:
                txt = self.text(b[0])[b[1]:e[1] + e[2]].decode()
                self.edits.replace(b[0], b[1], e[1] + e[2], "{" + ", ".join([txt] * self.N) + "}")

    # -- coalescing rewrite (converter.cpp:184-239, 275-330)
    def unconditional_calls(self, fn, names):
        """Calls to `names` in fn's body that are under no branch or loop."""
        found = []

        def visit(n, guarded):
            if not isinstance(n, dict):
                return
            g = guarded or n.get("kind") in LOOP_OR_BRANCH
            if n.get("kind") == "CallExpr" and not guarded:
                callee = self.callee_name(n)
                if callee in names:
                    found.append((n, callee))
            for c in n.get("inner", []):
                visit(c, g)
        for c in fn.get("inner", []):
            if c.get("kind") == "CompoundStmt":
                visit(c, False)
        return found

    @staticmethod
    def callee_name(call):
        def first_ref(n):
            if n.get("kind") == "DeclRefExpr":
                return n.get("referencedDecl", {}).get("name")
            for c in n.get("inner", [])[:1]:
                r = first_ref(c)
                if r:
                    return r
            return None
        inner = call.get("inner", [])
        return first_ref(inner[0]) if inner else None

    def add_arg(self, call, text):
        rparen = file_loc(call["_end"])
        if rparen is None:
            return
        nargs = len(call.get("inner", [])) - 1
        self.edits.insert(rparen, (", " if nargs > 0 else "") + text, rank=3, key="arg")

    def coalesce(self, ast):
        defs = {}
        for name, decls in self.funcs.items():
            for d in decls:
                if any(c.get("kind") == "CompoundStmt" for c in d.get("inner", [])):
                    defs[name] = d
        direct = {name for name, d in defs.items() if name != "main" and
                  self.unconditional_calls(d, {"malloc", "MPI_Alltoall"})}
        for name, d in defs.items():
            is_main = name == "main"
            arg = "true" if is_main else "__coalesced"
            if name in direct:
                nl = file_loc(d["_loc"])
                lp, rp = self.parens(nl[0], nl[1] + nl[2])
                params = [c for c in d.get("inner", []) if c.get("kind") == "ParmVarDecl"]
                inside = self.text(nl[0])[lp + 1:rp].decode().strip()
                if not params:
                    self.edits.replace(nl[0], lp + 1, rp, "bool __coalesced = false")
                elif inside:
                    self.edits.insert((nl[0], rp), ", bool __coalesced = false", rank=3, key="param")
            for call, callee in self.unconditional_calls(d, {"malloc", "MPI_Alltoall"} | direct):
                if callee == "MPI_Alltoall":
                    cl = file_loc(call["inner"][0]["_begin"])
                    ce = file_loc(call["inner"][0]["_end"])
                    self.edits.replace(cl[0], cl[1], ce[1] + ce[2], "MPI_Alltoall_coalesce")
                if callee in ("malloc", "MPI_Alltoall") or callee in direct:
                    if callee in direct and callee != name and False:
                        pass
                    self.add_arg(call, arg)
            if is_main and self.malloc_version in (3, 7):
                body = next(c for c in d["inner"] if c.get("kind") == "CompoundStmt")
                lb = file_loc(body["_begin"])
                self.edits.insert((lb[0], lb[1] + 1), "\n    __gpu_init_malloc();", rank=8, key="init")
                stmts = body.get("inner", [])
                if stmts:
                    last = file_loc(stmts[-1]["_begin"])
                    self.edits.insert(last, "__gpu_clean_malloc();\n    ", rank=0, key="clean")


def convert_file(path, project_dir=None, malloc_version=0, clang_args=(), ranks=None):
    """Returns {real path of each rewritten project file: converted bytes}."""
    project_dir = project_dir or os.environ.get("GPU_MPI_PROJECT") or os.path.dirname(os.path.abspath(path))
    ast = parse(path, clang_args)
    return Converter(path, project_dir, malloc_version, ranks).convert(ast)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Convert an MPI C program to gpumpi device code")
    ap.add_argument("source")
    ap.add_argument("--write-to-stdout", action="store_true")
    ap.add_argument("--malloc-version", type=int, default=0)
    ap.add_argument("--project", default=None, help="project directory (GPU_MPI_PROJECT)")
    ap.add_argument("-o", "--output", default=None, help="converted main file (default: source with .hip)")
    ap.add_argument("clang_args", nargs="*", help="arguments for clang after --")
    a = ap.parse_args(argv)
    out = convert_file(a.source, a.project, a.malloc_version, a.clang_args)
    main_real = os.path.realpath(a.source)
    for real, data in out.items():
        if a.write_to_stdout:
            if real == main_real:
                sys.stdout.write(data.decode())
            continue
        if real == main_real:
            dest = a.output or os.path.splitext(a.source)[0] + ".hip"
        else:
            dest = real + ".hip"
        with open(dest, "wb") as fh:
            fh.write(data)
    return 0


if __name__ == "__main__":
    sys.exit(main())
