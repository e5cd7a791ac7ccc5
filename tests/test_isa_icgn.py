"""Code generation of the IC-GN fast kernel (CPU; reads the gfx950 code object
of the library's own build, _build/icgn.o).

The instances that carry the 65 px class must keep the gather's LDS reads in
flight and fit their launch bound.  A source or compiler change can give
the unrolled slot loop an order in which every tap read waits for the one
before it (all reads into one register pair): the results do not change,
an evaluation takes 40% longer.  Only the generated code shows it."""
import collections
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "3d-stereo-aldic_amd", "_build", "icgn.o")
LLVM = "/opt/rocm/llvm/bin"
TOOLS = [os.path.join(LLVM, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump",
                                          "llvm-readelf")]

pytestmark = pytest.mark.skipif(not (os.path.exists(OBJ) and all(map(os.path.exists, TOOLS))),
                                reason="needs the built icgn.o and the ROCm LLVM tools")

# the IC-GN instances of the 65 px class (FS 65 or 0, list walk or not,
# reference interpolant or B-spline) and the persistent launch
C65 = "_ZN12_GLOBAL__N_111fast_kernelILi0ELi4ELi17ELi84ELi"
ONE_ENTRY = ("_ZN12_GLOBAL__N_111fast_kernelILi0ELi4ELi17ELi84ELi65ELb0ELi0EEEv9SolveArgs",
             "_ZN12_GLOBAL__N_111fast_kernelILi0ELi4ELi17ELi84ELi65ELb0ELi1EEEv9SolveArgs")
PERSIST = "_ZN12_GLOBAL__N_117persistent_kernelILi0E"


@pytest.fixture(scope="module")
def code(tmp_path_factory):
    d = tmp_path_factory.mktemp("isa")
    fat, co = str(d / "fat.bin"), str(d / "icgn.co")
    run = lambda *a: subprocess.run(a, check=True, capture_output=True, text=True).stdout
    run(TOOLS[0], f"--dump-section=.hip_fatbin={fat}", OBJ, str(d / "rest.o"))
    target = [t for t in run(TOOLS[1], "--list", "--type=o", f"--input={fat}").split()
              if t.endswith("gfx950")][0]
    run(TOOLS[1], "--unbundle", "--type=o", f"--input={fat}", f"--targets={target}",
        f"--output={co}")
    dis = run(TOOLS[2], "-d", "--symbolize-operands", co)
    notes = run(TOOLS[3], "--notes", co)
    # kernels -> basic blocks -> instruction lines
    funcs, cur, blk = {}, None, None
    for line in dis.splitlines():
        m = re.match(r"^[0-9a-f]+ <(\S+)>:$", line)
        if m:
            name = m.group(1)
            if name.startswith("L") and cur is not None:
                blk = name
                funcs[cur].setdefault(blk, [])
            else:
                cur, blk = name, "entry"
                funcs[cur] = {blk: []}
            continue
        if cur is not None and line.startswith("\t"):
            funcs[cur][blk].append(line.strip())
    # per kernel metadata: .name comes before .private_segment_fixed_size and
    # .vgpr_count in each (alphabetical) kernel map
    meta, name = {}, None
    for line in notes.splitlines():
        m = re.match(r"\s*\.name:\s+(\S+)", line)
        if m:
            name = m.group(1)
            meta[name] = {}
            continue
        m = re.match(r"\s*\.(private_segment_fixed_size|vgpr_count):\s+(\d+)", line)
        if m and name:
            meta[name][m.group(1)] = int(m.group(2))
    return funcs, meta


def gather_reads(blocks):
    """The two-tap LDS reads of the block with the most of them, by
    destination register pair."""
    best = max(blocks.values(), key=lambda b: sum(1 for i in b if i.startswith("ds_read2_b32")))
    return [i.split()[1].rstrip(",") for i in best if i.startswith("ds_read2_b32")]


def test_gather_reads_stay_in_flight(code):
    funcs, _ = code
    names = [k for k in funcs if k.startswith(C65) or k.startswith(PERSIST)]
    assert any(k in funcs for k in ONE_ENTRY) and any(k.startswith(PERSIST) for k in names)
    for k in names:
        reads = gather_reads(funcs[k])
        # 17 slots of 8 two-tap reads (16 full slots when the size is not fixed)
        assert len(reads) >= 128, (k, len(reads))
        top = collections.Counter(reads).most_common(1)[0][1] / len(reads)
        assert top <= 0.3, f"{k}: {top:.2f} of the gather's reads into one register pair"


def test_one_entry_instances_fit_three_waves(code):
    _, meta = code
    for k in ONE_ENTRY:
        assert meta[k]["vgpr_count"] <= 168, (k, meta[k])  # three waves per SIMD
        assert meta[k]["private_segment_fixed_size"] == 0, (k, meta[k])  # no spills
