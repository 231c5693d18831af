// Links libaqua.so: from $AQUA_LIB_DIR, or else from the package directory
// that `make -C aqua-engine_amd/csrc` builds it into (two levels up).
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=AQUA_LIB_DIR");
    let dir = match env::var("AQUA_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR"));
            manifest.join("..").join("..")
        }
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=aqua");
    // the library loads its HIP runtime from ROCm
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
}
