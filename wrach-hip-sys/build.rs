//! Links libwrach_hip.so.  WRACH_HIP_LIB_DIR names the directory that holds
//! it (built with `python -m wrach_amd.build`); the default is the
//! repository's `wrach_amd/` next to this crate.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = env::var("WRACH_HIP_LIB_DIR").map(PathBuf::from).unwrap_or_else(|_| {
        PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR"))
            .join("..")
            .join("wrach_amd")
    });
    println!("cargo:rerun-if-env-changed=WRACH_HIP_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=wrach_hip");
    // The library's own runtime dependency (ROCm's HIP runtime).
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
