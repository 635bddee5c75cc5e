"""The Rust binding crates (crates/sfk-sys, crates/sfk) against the C header,
on the CPU (no Rust toolchain is needed): the committed raw bindings are the
generator's output; every function of include/sfk.h, as gcc itself parses it
(-aux-info), is declared with the same parameter and return types; every
#[repr(C)] struct has the C struct's size and field offsets (gcc-compiled
offsetof); and the safe layer calls only declared functions."""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sfk.h")
SYS_RS = os.path.join(ROOT, "crates", "sfk-sys", "src", "lib.rs")
SAFE_RS = os.path.join(ROOT, "crates", "sfk", "src", "lib.rs")

C_TO_RUST = {"void": "c_void", "char": "c_char", "int": "c_int", "uint8_t": "u8",
             "int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64", "float": "f32",
             "double": "f64", "size_t": "usize"}


def gen_module():
    sys.path.insert(0, os.path.join(ROOT, "crates", "sfk-sys"))
    try:
        import gen_sys
        return gen_sys
    finally:
        sys.path.pop(0)


def c_type_to_rust(t):
    """gcc's spelling ('const float *', 'sfk_ctx **') -> the Rust type."""
    t = " ".join(t.replace("*", " * ").split())
    stars = t.count("*")
    words = [w for w in t.replace("*", " ").split() if w != "const"]
    assert len(words) == 1, t
    base = C_TO_RUST.get(words[0], words[0])
    if not stars:
        return base
    head = t.split("*")[0]
    r = ("*const " if "const" in head.split() else "*mut ") + base
    return "*mut " * (stars - 1) + r


def gcc_prototypes():
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "aux.txt")
        subprocess.run(["gcc", "-aux-info", out, "-fsyntax-only", "-x", "c", HEADER], check=True,
                       cwd=d)
        lines = open(out).read().splitlines()
    protos = {}
    for ln in lines:
        m = re.search(r"extern (.+?)\b(sfk_\w+) \((.*)\);", ln)
        if not m:
            continue
        ret = m.group(1).strip()
        args = [a.strip() for a in m.group(3).split(",")] if m.group(3).strip() != "void" else []
        protos[m.group(2)] = (ret, args)
    return protos


def rust_externs(text):
    body = text[text.index('unsafe extern "C" {'):]
    fns = {}
    for m in re.finditer(r"pub fn (sfk_\w+)\((.*?)\)( -> ([^;]+))?;", body, re.S):
        params = [p.strip() for p in m.group(2).split(",") if p.strip()]
        types = [p.split(":", 1)[1].strip() for p in params]
        fns[m.group(1)] = (m.group(4).strip() if m.group(4) else "c_void", types)
    return fns


def test_committed_bindings_are_the_generator_output():
    with open(SYS_RS) as f:
        assert f.read() == gen_module().generate(), \
            "crates/sfk-sys/src/lib.rs is stale: run python crates/sfk-sys/gen_sys.py"


def test_every_header_function_declared_with_its_types():
    protos = gcc_prototypes()
    fns = rust_externs(open(SYS_RS).read())
    assert len(protos) >= 78
    assert set(protos) == set(fns)
    for name, (ret, args) in protos.items():
        rret, rargs = fns[name]
        assert rret == c_type_to_rust(ret), (name, ret, rret)
        assert [c_type_to_rust(a) for a in args] == rargs, (name, args, rargs)


def rust_structs(text):
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\n#\[derive\(Clone, Copy, Debug\)\]\npub struct (\w+) \{\n(.*?)\n\}",
                         text, re.S):
        fields = []
        for ln in m.group(2).splitlines():
            fm = re.match(r"\s*pub (\w+): (.+),$", ln)
            fields.append((fm.group(1), fm.group(2)))
        out[m.group(1)] = fields
    return out


SIZES = {"u8": 1, "i32": 4, "u32": 4, "f32": 4, "c_int": 4, "u64": 8, "f64": 8, "usize": 8}


def rust_layout(fields, consts):
    """repr(C) offsets and size from the Rust field types."""
    off, align, offs = 0, 1, {}
    for name, t in fields:
        am = re.fullmatch(r"\[(.+); (.+)\]", t)
        if am:
            elem, count = am.group(1), am.group(2)
            count = int(consts[count.split()[0]]) if not count.isdigit() else int(count)
        else:
            elem, count = t, 1
        sz = 8 if elem.startswith("*") else SIZES[elem]
        off = (off + sz - 1) // sz * sz
        offs[name] = off
        off += sz * count
        align = max(align, sz)
    return offs, (off + align - 1) // align * align


def test_repr_c_structs_match_the_c_layout(tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    text = open(SYS_RS).read()
    consts = {m.group(1): m.group(2).rstrip("u32") for m in
              re.finditer(r"pub const (\w+): u32 = (\d+)u32;", text)}
    structs = rust_structs(text)
    assert {"sfk_camera", "sfk_view", "sfk_tile_isects", "sfk_render_buffers",
            "sfk_ply_layout", "sfk_sog_meta"} <= set(structs)
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for s, fields in structs.items():
        src.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f, _ in fields:
            src.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in
               subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    for s, fields in structs.items():
        offs, size = rust_layout(fields, consts)
        assert int(got[f"{s} size"]) == size, s
        for f, _ in fields:
            assert int(got[f"{s}.{f}"]) == offs[f], (s, f)


def test_safe_layer_calls_only_declared_functions():
    declared = set(rust_externs(open(SYS_RS).read()))
    safe = open(SAFE_RS).read()
    used = set(re.findall(r"sys::(sfk_\w+)\(", safe))
    assert used and used <= declared, used - declared
    # the reference's entry points the safe layer mirrors
    for name in ("fn prepare_isects", "fn finalize", "fn render_with", "fn argsort",
                 "fn render_responsibility", "fn accumulate_prepared", "fn new_uniform",
                 "fn eig", "fn update", "fn map_labels", "fn bits_for", "fn fractional_bits",
                 "fn from_bytes"):
        assert name in safe, name
    code = re.sub(r"//[^\n]*", "", safe)
    code = re.sub(r'"(\\.|[^"\\])*"', '""', code)
    for a, b in ("()", "{}", "[]"):
        assert code.count(a) == code.count(b), a
