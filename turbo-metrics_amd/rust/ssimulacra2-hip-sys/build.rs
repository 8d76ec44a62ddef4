// Links libssim2hip.so (built with `make -C turbo-metrics_amd`). Set
// SSIM2HIP_LIB_DIR to the folder that holds it; the HIP runtime it depends on
// (libamdhip64.so) is found through the usual loader paths.
fn main() {
    let dir = std::env::var("SSIM2HIP_LIB_DIR")
        .expect("set SSIM2HIP_LIB_DIR to the folder holding libssim2hip.so");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=ssim2hip");
    println!("cargo:rerun-if-env-changed=SSIM2HIP_LIB_DIR");
}
