"""Build-time check of the LDS a kernel was compiled with (run by the
Makefile after the library links).

Reads the gfx950 code object inside an object file built here (the
.hip_fatbin section, unbundled with clang-offload-bundler) and the kernel
descriptors' metadata (llvm-readelf --notes), and fails unless every kernel
whose (mangled) name matches the regular expression NAME has
.group_segment_fixed_size == BYTES.  The 3-D
advection adjoint declares exactly its LDS windows; any more (the compiler
placing private arrays in LDS beside them) has given wrong, run-to-run
varying gradients.

usage: python3 check_lds.py OBJECT NAME BYTES [--table OUT]
  --table OUT: also write every kernel's LDS, scratch, VGPR and spill counts.
"""
import os
import re
import subprocess
import sys
import tempfile

LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def kernels(obj):
    with tempfile.TemporaryDirectory() as d:
        fat = os.path.join(d, "fatbin")
        co = os.path.join(d, "co")
        subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fat}", obj],
                       check=True)
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", f"--input={fat}",
                        "--type=o", f"--targets={TARGET}", f"--output={co}"], check=True)
        notes = subprocess.run([f"{LLVM}/llvm-readelf", "--notes", co], check=True,
                               capture_output=True, text=True).stdout
    out = []
    cur = {}
    for line in notes.splitlines():
        m = re.match(r"\s*-?\s*\.(\w+):\s+(\S+)", line)
        if not m:
            continue
        key, val = m.group(1), m.group(2)
        if key in ("group_segment_fixed_size", "private_segment_fixed_size", "vgpr_count",
                   "vgpr_spill_count", "name"):
            cur[key] = val
        # a kernel's keys come in alphabetical order; .wavefront_size is its last
        if key == "wavefront_size" and "name" in cur:
            out.append(cur)
            cur = {}
    return out


def main(argv):
    obj, name, want = argv[1], argv[2], int(argv[3])
    ks = kernels(obj)
    hits = [k for k in ks if re.search(name, k["name"])]
    if not hits:
        print(f"check_lds: no kernel named *{name}* in {obj}", file=sys.stderr)
        return 1
    bad = [k for k in hits if int(k["group_segment_fixed_size"]) != want]
    for k in bad:
        print(f"check_lds: {k['name']} uses {k['group_segment_fixed_size']} B of LDS, "
              f"expected {want}", file=sys.stderr)
    if "--table" in argv:
        path = argv[argv.index("--table") + 1]
        with open(path, "w") as f:
            f.write("kernel\tlds_bytes\tscratch_bytes_per_lane\tvgprs\tvgpr_spills\n")
            for k in ks:
                f.write(f"{k['name']}\t{k['group_segment_fixed_size']}\t"
                        f"{k.get('private_segment_fixed_size', '?')}\t{k.get('vgpr_count', '?')}\t"
                        f"{k.get('vgpr_spill_count', '?')}\n")
    if bad:
        return 1
    print(f"check_lds: {len(hits)} {name} kernels at {want} B of LDS")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
