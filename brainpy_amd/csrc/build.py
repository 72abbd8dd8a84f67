"""Build libbpx.so (gfx950) in-tree with hipcc.

    python -m brainpy_amd.csrc.build        # or brainpy_amd.csrc.build.build()

Flags:
  --offload-arch=gfx950         CDNA4 only
  -munsafe-fp-atomics           hardware global_atomic_add_f32/f64
  -ffp-contract=off             one rounding per float operation (the neuron
                                update and jitconn weights are bit-exact with
                                their CPU restatement only without contraction)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libbpx.so")
PROBE_LIB = os.path.join(PKG, "libbpx_probe.so")   # measurement probes, loaded by the benchmark only
PROBE_SOURCES = ["bpx_probe.hip"]
SOURCES = ["bpx_runtime.hip", "bpx_csr.hip", "bpx_info.hip", "bpx_jitconn.hip", "bpx_neuron.hip", "bpx_tiled.hip",
           "bpx_net.hip", "bpx_plast.hip", "bpx_coo.hip", "bpx_conn.hip", "bpx_bsr.hip", "bpx_xla.hip",
           "bpx_batch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def _link(target, objs):
    tmp = target + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, target)


def build(verbose=False, jobs=8):
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES + PROBE_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    _link(LIB, objs[:len(SOURCES)])
    _link(PROBE_LIB, objs[len(SOURCES):])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
