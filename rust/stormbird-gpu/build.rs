//! Builds libstormbird_amd.so for gfx950 and links this crate against it.
//!
//! The library's own build (stormbird_amd/build.py: hipcc with -ffp-contract=off and packed
//! fp32 instructions off) is the one source of the compile flags, so it is run as is.
//! STORMBIRD_AMD_ROOT names the repository that holds stormbird_amd/ and include/ (default:
//! two directories above this crate); PYTHON the interpreter (default python3).  Follows the
//! build-script convention of interfaces/cstormbird/build.rs:6-21 (paths from
//! CARGO_MANIFEST_DIR, rerun when the sources change).

use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    let crate_dir = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let root = env::var("STORMBIRD_AMD_ROOT")
        .map(PathBuf::from)
        .unwrap_or_else(|_| crate_dir.join("..").join(".."));
    let python = env::var("PYTHON").unwrap_or_else(|_| "python3".to_string());

    let status = Command::new(&python)
        .args(["-m", "stormbird_amd.build"])
        .current_dir(&root)
        .status()
        .unwrap_or_else(|e| panic!("could not run `{python} -m stormbird_amd.build` in {}: {e}", root.display()));
    assert!(status.success(), "building libstormbird_amd.so failed ({status})");

    let lib_dir = root.join("stormbird_amd");
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=stormbird_amd");
    // tests and binaries of this package find the library without LD_LIBRARY_PATH
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
    println!("cargo:include={}", root.join("include").display());
    println!("cargo:rerun-if-changed={}", root.join("stormbird_amd").join("csrc").display());
    println!("cargo:rerun-if-changed={}", root.join("include").join("stormbird_gpu.h").display());
    println!("cargo:rerun-if-env-changed=STORMBIRD_AMD_ROOT");
}
