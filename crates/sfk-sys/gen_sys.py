"""Generate src/lib.rs, the raw Rust bindings of libsfk, from include/sfk.h.

Every #define, opaque handle, plain struct and function of the header
becomes a Rust constant, #[repr(C)] type or `extern "C"` declaration.  Run
`python crates/sfk-sys/gen_sys.py` after changing the header; the test suite
checks that the committed src/lib.rs is this script's output.
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "..", "include", "sfk.h")
OUT = os.path.join(HERE, "src", "lib.rs")

SCALARS = {
    "void": "c_void", "char": "c_char", "int": "c_int", "uint8_t": "u8", "int32_t": "i32",
    "uint32_t": "u32", "uint64_t": "u64", "float": "f32", "double": "f64", "size_t": "usize",
}
RUST_KEYWORDS = {"type", "ref", "in", "mod", "fn", "move", "box", "match", "self", "loop",
                 "use", "impl", "trait", "where", "crate", "super", "dyn", "async"}


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def rust_type(ctype, structs):
    """C type text (const, pointers) -> Rust type."""
    t = ctype.replace("const ", "const ").strip()
    stars = t.count("*")
    base = t.replace("*", " ").split()
    consts = [i for i, w in enumerate(base) if w == "const"]
    words = [w for w in base if w not in ("const", "struct")]
    assert len(words) == 1, ctype
    name = words[0]
    if name in SCALARS:
        r = SCALARS[name]
    elif name in structs:
        r = name
    else:
        raise ValueError(f"unknown C type {ctype!r}")
    if stars == 0:
        return r
    # Only the pointee of the innermost pointer may be const in this header
    # ("const T*", "const T**" = pointer to pointer to const T).
    inner_const = bool(consts) and consts[0] == 0
    r = ("*const " if inner_const else "*mut ") + r
    for _ in range(stars - 1):
        r = "*mut " + r
    return r


def split_decl(decl):
    """'const float* name[3]' -> ('const float*', 'name', '3' or None)."""
    decl = decl.strip()
    m = re.match(r"^(.*?)([A-Za-z_][A-Za-z0-9_]*)\s*(\[\s*([A-Za-z0-9_]+)\s*\])?$", decl)
    assert m, decl
    return m.group(1).strip(), m.group(2), m.group(4)


def parse(text):
    text = strip_comments(text)
    defines, opaque, structs, funcs = [], [], {}, []
    for m in re.finditer(r"^#define\s+(SFK_[A-Z0-9_]+)\s+(.+?)\s*$", text, re.M):
        if m.group(1) == "SFK_H":
            continue
        defines.append((m.group(1), m.group(2)))
    body = re.sub(r"^#.*$", " ", text, flags=re.M)
    body = body.replace('extern "C" {', " ").replace("}\n", "}\n")
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s+(\w+)\s*;", body):
        opaque.append(m.group(2))
    for m in re.finditer(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", body, re.S):
        fields = []
        for stmt in m.group(1).split(";"):
            stmt = " ".join(stmt.split())
            if not stmt:
                continue
            first, *rest = [p.strip() for p in stmt.split(",")]
            ctype, name, arr = split_decl(first)
            fields.append((ctype, name, arr))
            for r in rest:
                stars = r.count("*")
                _, n2, a2 = split_decl(r.replace("*", ""))
                fields.append((ctype + "*" * stars, n2, a2))
        structs[m.group(2)] = fields
    body = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", " ", body, flags=re.S)
    body = re.sub(r"typedef\s+struct\s+\w+\s+\w+\s*;", " ", body)
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_\s\*]*?)\b(sfk_\w+)\s*\(([^)]*)\)\s*;", body):
        ret = " ".join(m.group(1).split())
        params = []
        ptxt = " ".join(m.group(3).split())
        if ptxt and ptxt != "void":
            for p in ptxt.split(","):
                params.append(split_decl(p))
        funcs.append((ret, m.group(2), params))
    return defines, opaque, structs, funcs


def rust_const(name, value):
    v = value.strip()
    if v.startswith("(") and v.endswith(")"):
        v = v[1:-1].strip()
    v = re.sub(r"(\d+)u\b", r"\1u32", v)
    if name == "SFK_OK" or name.startswith("SFK_E_"):
        return f"pub const {name}: c_int = {v.replace('u32', '')};"
    if re.fullmatch(r"\d+", v):
        v = v + "u32" if not v.endswith("u32") else v
    return f"pub const {name}: u32 = {v};"


def generate(header=HEADER):
    with open(header) as f:
        defines, opaque, structs, funcs = parse(f.read())
    names = set(opaque) | set(structs)
    out = [
        "//! Raw bindings of libsfk, the MI355X (gfx950) render and B3-Seg evidence",
        "//! library (include/sfk.h).  Generated by gen_sys.py from the header; do not",
        "//! edit by hand.  Every call returns SFK_OK (0) or a negative SFK_E_* code,",
        "//! with sfk_last_error() holding the message; see the header for the",
        "//! contract of each entry point.",
        "#![allow(non_camel_case_types, non_snake_case, clippy::missing_safety_doc)]",
        "#![no_std]",
        "",
        "pub use core::ffi::{c_char, c_int, c_void};",
        "",
    ]
    for name, value in defines:
        out.append(rust_const(name, value))
    out.append("")
    for name in opaque:
        out += [f"/// Opaque handle (`{name}` of include/sfk.h).", "#[repr(C)]",
                f"pub struct {name} {{", "    _private: [u8; 0],", "}", ""]
    for name, fields in structs.items():
        out += ["#[repr(C)]", "#[derive(Clone, Copy, Debug)]", f"pub struct {name} {{"]
        for ctype, fname, arr in fields:
            rt = rust_type(ctype, names)
            if arr:
                rt = f"[{rt}; {arr if arr.isdigit() else arr + ' as usize'}]"
            out.append(f"    pub {fname}: {rt},")
        out += ["}", ""]
    out.append('unsafe extern "C" {')
    for ret, fname, params in funcs:
        args = []
        for ctype, pname, arr in params:
            rt = rust_type(ctype + ("*" if arr else ""), names)
            pn = f"r#{pname}" if pname in RUST_KEYWORDS else pname
            args.append(f"{pn}: {rt}")
        rr = rust_type(ret, names)
        tail = "" if rr == "c_void" else f" -> {rr}"
        line = f"    pub fn {fname}({', '.join(args)}){tail};"
        if len(line) > 100:
            inner = ",\n        ".join(args)
            line = f"    pub fn {fname}(\n        {inner},\n    ){tail};"
        out.append(line)
    out.append("}")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    text = generate()
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        with open(OUT) as f:
            sys.exit(0 if f.read() == text else 1)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        f.write(text)
    print(f"wrote {OUT}")
