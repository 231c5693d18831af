"""The Rust binding (aqua-engine_amd/rust/aqua-sys) against the C header it
mirrors. No Rust toolchain exists in this image, so the crate cannot be
compiled here; instead every `#[repr(C)]` struct and `extern "C"` function
of src/lib.rs is parsed and compared with include/aqua/aqua.h: the same
functions, the same parameter and return types (C -> Rust type map below),
the same struct fields in the same order, and the struct sizes the header
states."""
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "aqua", "aqua.h")
CRATE = os.path.join(ROOT, "aqua-engine_amd", "rust", "aqua-sys")

SCALAR = {"uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "uint8_t": "u8", "int": "c_int",
          "float": "f32", "double": "f64", "size_t": "usize", "char": "c_char", "void": "c_void"}
SIZE = {"u32": 4, "i32": 4, "u64": 8, "u8": 1, "c_int": 4, "f32": 4, "f64": 8, "usize": 8}


def _c_type(t):
    """C declaration type -> Rust type (pointers to opaque structs keep their name)."""
    t = t.strip()
    stars = t.count("*")
    base = t.replace("*", " ").split()
    const_first = base[0] == "const"
    base = [b for b in base if b != "const"]
    name = SCALAR.get(base[0], base[0])
    if stars == 0:
        return name
    if stars == 1:
        return ("*const " if const_first else "*mut ") + name
    assert stars == 2, t
    return "*mut " + ("*const " if const_first else "*mut ") + name


def _strip_c(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def header_functions():
    h = _strip_c(open(HEADER).read())
    out = {}
    for m in re.finditer(r"([\w\s\*]+?)\b(aqua_\w+)\s*\(([^)]*)\);", h):
        ret = " ".join(m.group(1).split())
        args = [] if m.group(3).strip() in ("", "void") else [a.strip() for a in m.group(3).split(",")]
        params = []
        for a in args:
            mm = re.match(r"(.*?)(\w+)$", a)
            params.append((mm.group(2), _c_type(mm.group(1))))
        out[m.group(2)] = (_c_type(ret), params)
    return out


def header_structs():
    h = _strip_c(open(HEADER).read())
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", h, re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            mm = re.match(r"(\w+) (.+)$", decl)
            ty = SCALAR[mm.group(1)]
            for name in mm.group(2).split(","):
                name = name.strip()
                arr = re.match(r"(\w+)\[(\d+)\]", name)
                fields.append((arr.group(1), f"[{ty}; {arr.group(2)}]") if arr else (name, ty))
        out[m.group(1)] = fields
    sizes = {}
    for m in re.finditer(r"/\* (\d+) bytes\.[^*]*\*/\s*typedef struct (\w+)", open(HEADER).read()):
        sizes[m.group(2)] = int(m.group(1))
    return out, sizes


def rust_ffi():
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    ffi = src[src.index("pub mod ffi {"):src.index("\nuse std::ffi")]
    structs = {}
    for m in re.finditer(r"pub struct (\w+) \{(.*?)\}", ffi, re.S):
        fields = []
        for line in m.group(2).split(","):
            line = line.strip()
            if not line:
                continue
            mm = re.match(r"(?:pub )?(\w+): (.+)$", line)
            fields.append((mm.group(1), mm.group(2).strip()))
        structs[m.group(1)] = fields
    fns = {}
    ext = ffi[ffi.index('extern "C" {'):]
    for m in re.finditer(r"pub fn (\w+)\(([^)]*)\)(?: -> ([^;]+))?;", ext, re.S):
        params = []
        for a in m.group(2).split(","):
            a = " ".join(a.split())
            if a:
                n, t = a.split(": ", 1)
                params.append((n, t))
        fns[m.group(1)] = ((m.group(3) or "()").strip(), params)
    return structs, fns


def test_every_function_is_bound_with_its_types():
    h = header_functions()
    _, r = rust_ffi()
    assert len(h) > 50
    assert sorted(h) == sorted(r), (set(h) ^ set(r))
    for name, (ret, params) in h.items():
        rret, rparams = r[name]
        want_ret = "()" if ret == "c_void" else ret
        assert rret == want_ret, (name, rret, want_ret)
        assert len(rparams) == len(params), name
        for (pn, pt), (rn, rt) in zip(params, rparams):
            assert rt == pt, (name, pn, rt, pt)


def test_every_struct_is_mirrored_field_by_field():
    hs, sizes = header_structs()
    rs, _ = rust_ffi()
    assert len(hs) >= 11
    for name, fields in hs.items():
        assert name in rs, name
        got = [(("type" if n == "type_" else n), t) for n, t in rs[name]]
        assert got == fields, (name, got, fields)
        # the size the header states, from the field types (no padding: the
        # header keeps 8-byte fields 8-byte aligned)
        size, off = 0, 0
        for _, t in fields:
            arr = re.match(r"\[(\w+); (\d+)\]", t)
            sz, al = (SIZE[arr.group(1)] * int(arr.group(2)), SIZE[arr.group(1)]) if arr else (SIZE[t], SIZE[t])
            assert off % al == 0, (name, off, t)
            off += sz
        if name in sizes:
            assert off == sizes[name], (name, off, sizes[name])


def test_crate_files_and_link_setup():
    toml = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'links = "aqua"' in toml and 'edition = "2018"' in toml   # as the reference crate (Cargo.toml:5)
    build = open(os.path.join(CRATE, "build.rs")).read()
    assert "cargo:rustc-link-lib=dylib=aqua" in build and "AQUA_LIB_DIR" in build
    ex = open(os.path.join(CRATE, "examples", "render.rs")).read()
    for call in ("Scene::load", "integrator_load", "Renderer::new", "render(", "tonemap(", "write_png("):
        assert call in ex, call


@pytest.mark.parametrize("const", ["AQUA_OK", "AQUA_ERR_IO", "AQUA_ERR_UNSUPPORTED", "AQUA_NRC_FP8",
                                   "AQUA_PROCGEN_INTERIOR", "AQUA_INTEGRATOR_NRC", "AQUA_ABI_VERSION"])
def test_constants_match(const):
    h = open(HEADER).read()
    m = re.search(rf"\b{const}\b\s*=?\s*(\d+)", h)
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    r = re.search(rf"pub const {const}: \w+ = (\d+);", src)
    assert m and r and m.group(1) == r.group(1), const
