//! Links libgf2ldpc.so: from GF2LDPC_LIB_DIR, else from the HIP build's
//! package directory next to this crate (../../gf2_amd).
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=GF2LDPC_LIB_DIR");
    let dir = match env::var("GF2LDPC_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../gf2_amd"),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=gf2ldpc");
    // the shared library is found at run time through the same directory
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
