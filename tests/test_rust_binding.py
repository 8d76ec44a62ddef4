"""The Rust drop-in crates (turbo-metrics_amd/rust) cannot be built here (no
Rust toolchain); this checks that the -sys crate declares exactly the
functions of include/ssim2hip.h and the safe crate keeps the reference
type's method names (crates/ssimulacra2-cuda/src/lib.rs)."""
import re

from conftest import ROOT

RUST = ROOT / "turbo-metrics_amd" / "rust"


def test_sys_crate_declares_every_header_function():
    hdr = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "ssim2hip.h").read_text(), flags=re.S)
    header = set(re.findall(r"\b(ssim2_\w+)\s*\(", hdr))
    sys_rs = (RUST / "ssimulacra2-hip-sys" / "src" / "lib.rs").read_text()
    rust = set(re.findall(r"pub fn (ssim2_\w+)\s*\(", sys_rs))
    assert header == rust, (sorted(header - rust), sorted(rust - header))


def test_safe_crate_keeps_the_reference_methods():
    src = (RUST / "ssimulacra2-hip" / "src" / "lib.rs").read_text()
    methods = set(re.findall(r"pub fn (\w+)\s*\(", src))
    for m in ("new", "compute", "get_score", "compute_sync", "compute_srgb_sync", "compute_from_cpu_srgb_sync",
              "compute_from_cpu_dis_only_srgb_sync", "prepare_reference_from_cpu_srgb", "set_reference_linear",
              "compute_with_reference_linear", "clear_reference", "has_reference", "mem_usage", "kernel",
              "srgb_to_linear", "downscale_by_2", "linear_to_xyb"):
        assert m in methods, m
