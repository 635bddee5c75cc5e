//! Links libsfk.so, built by `python -m splatfield_amd.build` (hipcc for
//! gfx950) into splatfield_amd/ of this repository, or found in
//! $SFK_LIB_DIR.  At run time the loader must find it too (rpath below, or
//! LD_LIBRARY_PATH).
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("SFK_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap())
            .join("..")
            .join("..")
            .join("splatfield_amd"),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=sfk");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    println!("cargo:rerun-if-env-changed=SFK_LIB_DIR");
    println!("cargo:rerun-if-changed=../../include/sfk.h");
}
