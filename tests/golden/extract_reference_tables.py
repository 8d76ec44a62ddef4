"""Extract the reference-held constant tables of the ssimulacra2 CPU path into a
JSON fixture (data only: numbers, no source text).

Run in a container that has the reference checkout mounted at /root/reference;
the output `reference_tables.json` is committed and is what the tests read.

Tables taken (all from crates/ssimulacra2-cuda/examples/cpu.rs):
  * FROM_SRGB8_TABLE[256]            cpu.rs:20-277   (f32)
  * Msssim::score WEIGHT[108]        cpu.rs:729-838  (f64)
  * recursive-Gaussian consts        cpu.rs:931-948  (f32)
  * opsin/XYB scalar constants       cpu.rs:421-434  (f32 literals)
The same LUT also appears in crates/ssimulacra2-cuda-kernel/src/srgb.rs:18-51 and
the same WEIGHT table in crates/ssimulacra2-cuda/src/lib.rs:1023-1153; the script
checks that those copies agree before writing.
"""
import json
import re
import struct
import sys
from pathlib import Path

REF = Path("/root/reference")
CPU_RS = REF / "crates/ssimulacra2-cuda/examples/cpu.rs"
SRGB_RS = REF / "crates/ssimulacra2-cuda-kernel/src/srgb.rs"
LIB_RS = REF / "crates/ssimulacra2-cuda/src/lib.rs"

NUM = r"-?[0-9][0-9_]*\.?[0-9_]*(?:e-?[0-9]+)?"


def _numbers(block):
    out = []
    for tok in re.findall(NUM, block):
        out.append(float(tok.replace("_", "")))
    return out


def _array(src, header):
    i = src.index(header)
    j = src.index("];", i)
    body = src[i + len(header):j]
    body = re.sub(r"//[^\n]*", "", body)
    body = re.sub(r"_f(32|64)", "", body)
    return _numbers(body)


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def f64_bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main(out_path):
    cpu = CPU_RS.read_text()
    lut = _array(cpu, "FROM_SRGB8_TABLE: [f32; 256] = [")
    assert len(lut) == 256, len(lut)
    lut_k = _array(SRGB_RS.read_text(), "SRGB8_TO_LINEARF32_LUT: [f32; 256] = [")
    assert [f32_bits(v) for v in lut_k] == [f32_bits(v) for v in lut], "kernel LUT differs"

    w = _array(cpu, "const WEIGHT: [f64; 108] = [")
    assert len(w) == 108, len(w)
    w_lib = _array(LIB_RS.read_text(), "const WEIGHT: [f64; 108] = [")
    assert [f64_bits(v) for v in w_lib] == [f64_bits(v) for v in w], "lib.rs WEIGHT differs"

    consts = {}
    m = re.search(r"mod consts \{(.*?)\n\}", cpu, re.S)
    for name, val in re.findall(r"pub const (\w+): f32 = (" + NUM + r")_f32;", m.group(1)):
        consts[name] = float(val.replace("_", ""))
    rad = re.search(r"pub const RADIUS: usize = (\d+)_usize;", m.group(1))
    consts_radius = int(rad.group(1))

    xyb = {}
    for name in ["K_M02", "K_M00", "K_M12", "K_M10", "K_M20", "K_M21", "K_B0", "K_B0_ROOT"]:
        mm = re.search(r"const " + name + r": f32 = (" + NUM + r")(?:_?f32)?;", cpu)
        xyb[name] = float(mm.group(1).replace("_", ""))

    doc = {
        "source": "crates/ssimulacra2-cuda/examples/cpu.rs",
        "srgb8_lut_f32_bits": [f32_bits(v) for v in lut],
        "weight_f64_bits": [f64_bits(v) for v in w],
        "iir_radius": consts_radius,
        "iir_f32_bits": {k: f32_bits(v) for k, v in sorted(consts.items())},
        "xyb_f32_literals": xyb,
    }
    Path(out_path).write_text(json.dumps(doc, indent=1) + "\n")
    print("wrote", out_path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else str(Path(__file__).with_name("reference_tables.json")))
