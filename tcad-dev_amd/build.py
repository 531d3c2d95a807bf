"""Build libtcadhip.so for gfx950 in-tree (tcad-dev_amd/tcadhip/libtcadhip.so).

Every HIP translation unit is compiled with -ffp-contract=off (no fused
multiply-add, so the residual rounds like the reference's NumPy
expressions) and with packed-FP32 instructions disabled.

TCADHIP_BS=<threads> builds another node-workgroup size (default 128),
TCADHIP_PY=<rows> another 2D patch height (default 4), TCADHIP_T3=<x>x<y>
another 3D tile extent in x and y (default 2x1; the z extent makes up the
16 nodes of a tile), and TCADHIP_OUT=<path> writes the library elsewhere
(tcadhip loads such a build through TCADHIP_LIB).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.environ.get("TCADHIP_OUT") or os.path.join(HERE, "tcadhip", "libtcadhip.so")
BS = os.environ.get("TCADHIP_BS")          # node workgroup size (default 128)
PY = os.environ.get("TCADHIP_PY")          # patch height (default 4)
T3 = os.environ.get("TCADHIP_T3")          # 3D tile x and y extent (default "2x1")
OBJ = os.path.join(HERE, "_build" + (f"_bs{BS}" if BS else "") + (f"_py{PY}" if PY else "")
                   + (f"_t3{T3}" if T3 else ""))
SOURCES = ["assemble.hip", "krylov.hip", "batch.hip", "ac.hip", "csr.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(REPO, "include"), "-I", CSRC, "-Wall", "-Wno-unused-result"]
if BS:
    FLAGS.append(f"-DTCAD_BS={int(BS)}")
if PY:
    FLAGS.append(f"-DTCAD_PY={int(PY)}")
if T3:
    t3x, t3y = (int(v) for v in T3.split("x"))
    FLAGS += [f"-DTCAD_T3X={t3x}", f"-DTCAD_T3Y={t3y}"]


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    if r.stderr.strip():
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
