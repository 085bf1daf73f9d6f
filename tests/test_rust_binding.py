"""The Rust crate's raw binding (rust/stormbird-gpu/src/lib.rs, `mod ffi`) declares exactly
the functions of include/stormbird_gpu.h, with the same argument count and matching types,
and the sb_step_result struct field for field.  (No Rust toolchain is needed: both files are
read as text.)"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stormbird_gpu.h")
CRATE = os.path.join(ROOT, "rust", "stormbird-gpu")

C_TO_RUST = {
    "void": None, "const char *": "*const c_char", "char *": "*mut c_char", "long": "c_long",
    "long *": "*mut c_long", "const long *": "*const c_long", "size_t": "usize", "int": "c_int",
    "int *": "*mut c_int", "double": "f64", "double *": "*mut f64", "const double *": "*const f64",
    "float": "f32", "float *": "*mut f32", "const float *": "*const f32", "unsigned long long": "u64",
    "const char * const *": "*const *const c_char", "void *": "*mut c_void",
    "sb_step_result *": "*mut sb_step_result",
}
for _o in ("sb_sim", "sb_batch", "sb_mlp", "sb_csm"):
    C_TO_RUST[_o + " *"] = "*mut " + _o
    C_TO_RUST["const " + _o + " *"] = "*const " + _o


def _norm(t):
    return " ".join(t.replace("*", " * ").split())


def header_functions():
    s = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"^([A-Za-z_][\w \*]*?)\b(sb_\w+)\s*\(([^;]*?)\)\s*;", s, flags=re.M):
        a = " ".join(args.split())
        types = [] if a == "void" else [_norm(re.match(r"(.*?)\w+$", p.strip()).group(1)) for p in a.split(",")]
        out[name] = (C_TO_RUST[_norm(ret)], [C_TO_RUST[t] for t in types])
    return out


def rust_functions():
    s = open(os.path.join(CRATE, "src", "lib.rs")).read()
    block = s[s.index('extern "C" {'):]
    out = {}
    for name, args, ret in re.findall(r"pub fn (sb_\w+)\(([^)]*)\)(?:\s*->\s*([^;]+))?;", block):
        types = [a.split(":", 1)[1].strip() for a in args.split(",") if a.strip()]
        out[name] = (ret.strip() if ret else None, types)
    return out


def test_every_header_function_is_bound_with_its_signature():
    h, r = header_functions(), rust_functions()
    assert len(h) > 80
    assert set(h) == set(r), f"only in the header: {set(h) - set(r)}; only in Rust: {set(r) - set(h)}"
    for name in h:
        assert h[name] == r[name], f"{name}: header {h[name]} vs Rust {r[name]}"


def test_step_result_layout_matches():
    s = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct sb_step_result \{(.*?)\} sb_step_result;", s, flags=re.S).group(1)
    c_fields = [(_norm(m.group(1)), m.group(2)) for m in re.finditer(r"([\w \*]+?)\s*\b(\w+);", body)]
    rs = open(os.path.join(CRATE, "src", "lib.rs")).read()
    rbody = re.search(r"pub struct sb_step_result \{(.*?)\n    \}", rs, flags=re.S).group(1)
    r_fields = re.findall(r"pub (\w+): ([^,]+),", rbody)
    assert [n for _, n in c_fields] == [n for n, _ in r_fields]
    for (ct, n), (_, rt) in zip(c_fields, r_fields):
        assert C_TO_RUST[ct] == rt.strip(), n


def test_crate_files_present():
    for f in ("Cargo.toml", "build.rs", os.path.join("src", "lib.rs"), os.path.join("tests", "compare_with_cpu.rs")):
        assert os.path.exists(os.path.join(CRATE, f)), f
    build = open(os.path.join(CRATE, "build.rs")).read()
    assert "rustc-link-lib=dylib=stormbird_amd" in build and "stormbird_amd.build" in build
