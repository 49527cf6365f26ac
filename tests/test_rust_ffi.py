"""The Rust FFI crate (crates/gf2-ldpc-hip) against the C header it binds.

cargo is not available here, so the crate is checked as text: every
function of include/gf2ldpc.h is declared in src/ffi.rs's extern block with
the same name, the same number of parameters and the Rust counterpart of
each C type (pointer constness included), nothing else is declared there,
and every #define constant has a `pub const` of the same value.  The crate's
own GPU test (tests/decode_batch.rs) compares it with gf2-coding's
LdpcDecoder::decode_batch on a host with cargo and a HIP device.
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gf2ldpc.h")
CRATE = os.path.join(ROOT, "crates", "gf2-ldpc-hip")

BASE = {"int": "c_int", "void": "c_void", "char": "c_char", "double": "f64", "float": "f32", "size_t": "usize",
        "uint8_t": "u8", "uint16_t": "u16", "uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64",
        "gf2l_ctx": "gf2l_ctx", "gf2l_code": "gf2l_code", "gf2l_nms": "gf2l_nms", "gf2l_codeset": "gf2l_codeset",
        "gf2_ldpc_decoder": "gf2_ldpc_decoder"}


def strip_c(src):
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    return "\n".join(l for l in src.split("\n") if not l.strip().startswith("#"))


def c_functions(src):
    out = {}
    for st in strip_c(src).split(";"):
        st = " ".join(st.split())
        m = re.match(r'^(?:extern "C" \{ )?(.*?)\b(gf2\w*)\s*\((.*)\)$', st)
        if m and "typedef" not in st:
            args = [] if m.group(3).strip() in ("", "void") else [a.strip() for a in m.group(3).split(",")]
            out[m.group(2)] = (m.group(1).strip(), args)
    return out


def c_type_to_rust(ctype):
    """C type (without the parameter name) -> the Rust FFI type."""
    toks = re.findall(r"\w+|\*", ctype)
    base_const = False
    i = 0
    while toks[i] == "const":
        base_const = True
        i += 1
    base = BASE[toks[i]]
    i += 1
    if i < len(toks) and toks[i] == "const":  # "T const"
        base_const = True
        i += 1
    rust, pointee_const = base, base_const
    while i < len(toks):
        assert toks[i] == "*", ctype
        rust = ("*const " if pointee_const else "*mut ") + rust
        i += 1
        pointee_const = i < len(toks) and toks[i] == "const"
        if pointee_const:
            i += 1
    return "()" if rust == "c_void" else rust


def c_param_type(param):
    """'const float *llr' -> 'const float *' (drop the name)."""
    m = re.match(r"^(.*?)(\w+)$", param.strip())
    return m.group(1).strip()


def rust_functions(src):
    block = src[src.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (\w+)\((.*?)\)\s*(?:->\s*([^;]+?))?\s*;", block, flags=re.S):
        params = [p.strip() for p in " ".join(m.group(2).split()).split(",") if p.strip()]
        types = [p.split(":", 1)[1].strip() for p in params]
        out[m.group(1)] = ((m.group(3) or "()").strip(), types)
    return out


def mismatches(header_src, ffi_src):
    cf = c_functions(header_src)
    rf = rust_functions(ffi_src)
    bad = [("missing", n) for n in sorted(set(cf) - set(rf))] + [("extra", n) for n in sorted(set(rf) - set(cf))]
    for name in sorted(set(cf) & set(rf)):
        (ret, args), (rret, rargs) = cf[name], rf[name]
        if c_type_to_rust(ret) != rret or len(args) != len(rargs) or \
                any(c_type_to_rust(c_param_type(a)) != r for a, r in zip(args, rargs)):
            bad.append(("signature", name))
    return bad, len(cf)


def test_ffi_declares_every_header_function_with_matching_types():
    ffi = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    bad, count = mismatches(open(HEADER).read(), ffi)
    assert count >= 50 and not bad, bad
    # the check notices a changed constness, a dropped parameter and a missing item
    assert mismatches(open(HEADER).read(), ffi.replace("llr: *const f32,", "llr: *mut f32,", 1))[0]
    assert mismatches(open(HEADER).read(), ffi.replace("max_it: u32,\n", "", 1))[0]
    assert mismatches(open(HEADER).read(), ffi.replace("pub fn gf2l_stream", "fn gf2l_stream", 1))[0]


def define_values(src):
    out = {}
    for name, value, shift in re.findall(r"#define\s+(GF2L_\w+)\s+\(?(-?\d+)u?(?:\s*<<\s*(\d+))?\)?", src):
        out[name] = int(value) << int(shift) if shift else int(value)
    return out


def test_ffi_constants_match_header():
    defines = define_values(open(HEADER).read())
    rust = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    consts = {}
    for name, value, shift in re.findall(r"pub const (\w+): \w+ = (-?\d+)(?: << (\d+))?;", rust):
        consts[name] = int(value) << int(shift) if shift else int(value)
    assert len(defines) >= 20
    assert defines == consts, sorted(set(defines.items()) ^ set(consts.items()))


def test_type_mapping_examples():
    assert c_type_to_rust("const gf2l_code *const *") == "*const *const gf2l_code"
    assert c_type_to_rust("gf2l_ctx **") == "*mut *mut gf2l_ctx"
    assert c_type_to_rust("const void *") == "*const c_void"
    assert c_type_to_rust("uint8_t *") == "*mut u8"
    assert c_type_to_rust("size_t") == "usize"


def test_crate_layout():
    for rel in ["Cargo.toml", "build.rs", "src/lib.rs", "src/ffi.rs", "tests/decode_batch.rs"]:
        assert os.path.exists(os.path.join(CRATE, rel)), rel
    lib = open(os.path.join(CRATE, "src", "lib.rs")).read()
    # the drop-in of LdpcDecoder::decode_batch and the decoder traits
    assert re.search(r"pub fn decode_batch\(code: &HipCode, llr_blocks: &\[Vec<Llr>\], max_iterations: usize\)"
                     r" -> Vec<DecoderResult>", lib)
    assert "impl IterativeSoftDecoder for HipLdpcDecoder" in lib
    assert "impl SoftDecoder for HipLdpcDecoder" in lib
    cargo = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'links = "gf2ldpc"' in cargo and "gf2-coding" in cargo
