"""Resource use of the compiled pass kernels (no device needed).

A pass kernel may use scratch (private) memory only when the register
budget of its variant forces spills: the allocator then sits at that cap
(512 / W VGPRs for W waves per SIMD), and the autotuner measures the variant
against the others.  Scratch below the cap has a structural cause, such as
a call in the per-point loop, which makes every value live across it spill
even when the call never runs; that traffic goes to HBM: with one
out-of-line library call in the phase path the 256^3 complex64 Z_MID pass
wrote 28 MiB per launch beyond its 128 MiB of field.  The check reads the
kernel metadata of the gfx950 code objects that the build embeds in each
per-variant object file.
"""

import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BUILD = ROOT / "gpe_simulator_amd" / "_build"
LLVM = Path("/opt/rocm/llvm/bin")
TOOLS = {name: LLVM / name for name in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")}
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def kernel_metadata(obj, tmp):
    """(name, private_segment_fixed_size, vgpr_count) of every kernel in `obj`."""
    fat, code = tmp / (obj.stem + ".fatbin"), tmp / (obj.stem + ".co")
    subprocess.run([str(TOOLS["llvm-objcopy"]), f"--dump-section=.hip_fatbin={fat}", str(obj), str(tmp / "host.o")],
                   check=True, capture_output=True)
    subprocess.run([str(TOOLS["clang-offload-bundler"]), "--unbundle", "--type=o", f"--targets={TARGET}",
                    f"--input={fat}", f"--output={code}"], check=True, capture_output=True)
    notes = subprocess.run([str(TOOLS["llvm-readelf"]), "--notes", str(code)], check=True, capture_output=True,
                           text=True).stdout
    out = []
    for block in notes.split("  - .")[1:]:
        name = re.search(r"\.name:\s+(\S+)", "." + block)
        scratch = re.search(r"\.private_segment_fixed_size:\s+(\d+)", "." + block)
        vgpr = re.search(r"\.vgpr_count:\s+(\d+)", "." + block)
        if name and scratch:
            out.append((name.group(1), int(scratch.group(1)), int(vgpr.group(1)) if vgpr else None))
    return out


def current_variants():
    """{(precision, side, variant): waves per SIMD} of the variants the build
    compiles (the generic transform of each precision has no register cap)."""
    text = (ROOT / "gpe_simulator_amd" / "csrc" / "gpe_shapes.h").read_text()
    block = text[text.index("#define GPE_VARIANTS(X)"):text.index("namespace gpe")]
    out = {(0, 0, 0): None, (1, 0, 0): None}
    for p, n, v, _tp, _b, w, _rd in re.findall(
            r"X\((\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", block):
        out[(int(p), int(n), int(v))] = int(w)
    return out


def register_cap(waves):
    return 512 if waves is None else min(512, (512 // waves) // 8 * 8)


@pytest.mark.skipif(not all(p.exists() for p in TOOLS.values()), reason="ROCm LLVM tools not installed")
def test_pass_kernels_spill_only_at_their_register_cap(tmp_path):
    variants = current_variants()
    objects = {key: BUILD / f"gpe_pass_p{key[0]}_n{key[1]}_v{key[2]}.o" for key in variants}
    if not all(o.exists() for o in objects.values()):
        pytest.skip("library not built (python -m gpe_simulator_amd.build)")
    checked, structural = 0, []
    for key, obj in objects.items():
        cap = register_cap(variants[key])
        for name, scratch, vgpr in kernel_metadata(obj, tmp_path):
            if "k_pass" not in name:
                continue
            checked += 1
            if scratch and (vgpr is None or vgpr < cap - 8):
                structural.append(f"{obj.name}: {name} scratch {scratch} B/lane at {vgpr} of {cap} VGPRs")
    assert checked >= len(objects)
    assert not structural, "\n".join(structural)


@pytest.mark.skipif(not all(p.exists() for p in TOOLS.values()), reason="ROCm LLVM tools not installed")
def test_plane_kernels_spill_only_at_their_register_cap(tmp_path):
    """The plane passes (one block of n pencils x TP threads per x plane,
    compiled in each side's first variant) likewise: a block of T threads
    puts ceil(T / 256) waves on each SIMD, so the cap is 512 over that."""
    variants = current_variants()
    objects = [BUILD / f"gpe_pass_p{p}_n{n}_v0.o" for (p, n, v) in variants if v == 0 and n]
    if not all(o.exists() for o in objects):
        pytest.skip("library not built (python -m gpe_simulator_amd.build)")
    checked, structural = 0, []
    for obj in objects:
        for name, scratch, vgpr in kernel_metadata(obj, tmp_path):
            m = re.search(r"k_plane<(?:float|double), (\d+), (\d+), (\d+)>", name) or \
                re.search(r"k_planeI[fd]Li(\d+)ELi(\d+)ELi(\d+)E", name)
            if not m:
                continue
            n, tp = int(m.group(1)), int(m.group(2))
            waves_per_simd = -(-(n * tp) // 256)
            cap = min(512, (512 // waves_per_simd) // 8 * 8)
            checked += 1
            if scratch and (vgpr is None or vgpr < cap - 8):
                structural.append(f"{obj.name}: {name} scratch {scratch} B/lane at {vgpr} of {cap} VGPRs")
    assert checked >= 5 * 5      # five plane kinds at each of at least five (precision, side)
    assert not structural, "\n".join(structural)


def test_headline_kernels_use_no_scratch(tmp_path):
    """The 256^3 variants (both precisions) run without any scratch."""
    variants = current_variants()
    objects = [BUILD / f"gpe_pass_p{p}_n{n}_v{v}.o" for (p, n, v) in variants if n == 256]
    if not objects or not all(o.exists() for o in objects):
        pytest.skip("library not built (python -m gpe_simulator_amd.build)")
    if not all(p.exists() for p in TOOLS.values()):
        pytest.skip("ROCm LLVM tools not installed")
    for obj in objects:
        for name, scratch, vgpr in kernel_metadata(obj, tmp_path):
            if "k_pass" in name:
                assert scratch == 0, f"{obj.name}: {name} uses {scratch} B/lane of scratch ({vgpr} VGPRs)"


def test_variants_of_one_shape_share_their_arithmetic():
    """Every compiled variant of one (precision, side) has the same threads per
    pencil, so the same codelets, twiddles and summation order: the tuner's
    choice of tile cannot change a result, and a batch member stays
    bit-identical to the same solve stepped alone whatever each was tuned to."""
    text = (ROOT / "gpe_simulator_amd" / "csrc" / "gpe_shapes.h").read_text()
    block = text[text.index("#define GPE_VARIANTS(X)"):text.index("namespace gpe")]
    tps = {}
    for p, n, _v, tp, _b, _w, _rd in re.findall(
            r"X\((\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", block):
        tps.setdefault((int(p), int(n)), set()).add(int(tp))
    assert tps
    mixed = {shape: sorted(t) for shape, t in tps.items() if len(t) > 1}
    assert not mixed, f"shapes with variants of different threads per pencil: {mixed}"
