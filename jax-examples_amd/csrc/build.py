"""Build libqcmd.so for gfx950 from the HIP sources in this directory.

Each ERI class (and precision) is its own translation unit, generated into
``_build/gen``, so the 42 heavily unrolled kernels compile in parallel.
Usage: ``python build.py [--jobs N] [--reuse]``; ``--reuse`` keeps objects
whose sources (including headers) are unchanged.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
BUILD = os.path.join(PKG, "_build")
GEN = os.path.join(BUILD, "gen")
OBJ = os.path.join(BUILD, "obj")
LIB_DIR = os.path.join(PKG, "jax_qc", "_lib")
LIB = os.path.join(LIB_DIR, "libqcmd.so")

ERI_CLASSES = [
    (0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1),
    (2, 0, 0, 0), (2, 0, 1, 0), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 0, 0), (2, 1, 1, 0),
    (2, 1, 1, 1), (2, 1, 2, 0), (2, 1, 2, 1), (2, 2, 0, 0), (2, 2, 1, 0), (2, 2, 1, 1),
    (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2),
]
# derivative-ERI classes that also get the one-thread-per-quartet kernel
# (the rest have only the one-wave-per-quartet form: their G blocks and
# Hermite tables do not fit a thread's registers)
GRAD_PER_THREAD = {"0000", "1000", "1010", "1100", "1110", "2000", "2010", "2100", "2020"}
# classes whose one-thread-per-quartet kernel fits the registers in FP32 only
# (measured faster than the wave form on (H2O)_16: (dd|ss) 6.2 vs 9.2 ms;
# (ds|pp) keeps arrays on the stack, (dp|ps) and (dp|ds) hold one wave per
# SIMD at ~400 registers, and those lose: 760 vs 154 ms for (dp|ps))
GRAD_PER_THREAD_F32 = GRAD_PER_THREAD | {"2200"}
# derivative-ERI classes with the one-lane-per-primitive-quartet kernel
# (grad_prim.h, one launch per ket Cartesian pair): those whose R and h lists
# fit a lane's registers without spills (FP64 (ds|pp) 239 VGPRs, (pp|pp) 256;
# FP32 (ds|pp) 134, (dp|pp) 233, (dp|ds) 229; FP64 (dp|ps) and (dd|ss) and
# FP32 (dd|ps) spill 100-700 registers)
GRAD_PRIM = {"1111", "2011"}
# FP32 also (dp|ps), (dp|pp), (dp|ds) (the library default now takes the
# G-first form below for all of these; this form stays selectable as variant 3)
GRAD_PRIM_F32 = GRAD_PRIM | {"2110", "2111", "2120"}
# ket Cartesian pairs per launch of that kernel, (FP64, FP32), 1 where not
# listed.  Measured on the (H2O)_16 gradient (1 -> chosen): FP32 (ds|pp)
# 142.9 -> 21.8 ms at 9 (53.9 at 3), FP32 (pp|pp) 102.9 -> 39.4 ms at 3
# (47.3 at 9); larger chunks lost for FP64 (pp|pp) (444 -> 1144 ms at 3),
# FP64 (ds|pp) (694 -> 876 at 9) and FP32 (dp|ps), (dp|pp), (dp|ds)
# (105 -> 354, 92 -> 423, 41 -> 162 ms at 3)
GRAD_PRIM_CHUNK = {"2011": (1, 9), "1111": (1, 3)}
# derivative-ERI classes with the one-lane-per-primitive-quartet kernel that
# contracts G with the bra E first (variant 5, grad_prim.h grad2e_yprim_kernel)
# and their ket Cartesian pairs per launch (FP64, FP32): the LDS holds Y for
# 9 bra primitives x the launch's ket pairs
GRAD_YPRIM = {"1111", "2011", "2110", "2111", "2120", "2121", "2200", "2210", "2211", "2220",
              "2221"}
GRAD_YPRIM_F32 = GRAD_YPRIM - {"2200"}
GRAD_YPRIM_CHUNK = {"1111": (3, 9), "2011": (3, 9), "2110": (3, 3), "2111": (3, 3),
                    "2120": (3, 3), "2121": (3, 3), "2200": (1, 1), "2210": (1, 3),
                    "2211": (1, 3), "2220": (1, 3), "2221": (1, 3)}
# at most this many kernel instantiations (launch chunks) per translation unit
YPRIM_CHUNKS_PER_UNIT = 3
# derivative-ERI classes that also get the wave kernel with its Y phase on the
# matrix cores (variant 4; measured against the VALU Y phase)
GRAD_WAVE_MFMA = {"1111", "2011", "2110", "2111", "2121"}
STATIC_SOURCES = ["capi.hip", "oneint.hip", "jk.hip", "grad.hip", "purify.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-I" + HERE, "-I" + os.path.join(REPO, "include"),
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
] + os.environ.get("QCMD_EXTRA_FLAGS", "").split()  # e.g. -DQC_EW_KC_MIN=2 for tuning sweeps


def _generate():
    os.makedirs(GEN, exist_ok=True)
    out = []
    for c in ERI_CLASSES:
        tag = "".join(map(str, c))
        for prec, real in (("f64", "double"), ("f32", "float")):
            path = os.path.join(GEN, f"eri_{tag}_{prec}.hip")
            args = f"{c[0]}, {c[1]}, {c[2]}, {c[3]}, {real}"
            text = (
                '#include "eri_coop.h"\n'
                "namespace qcmd {\n"
                f"hipError_t eri_launch_{tag}_{prec}(const EriArgs& a, hipStream_t s) {{\n"
                f"  return eri_launch_class<{args}>(a, s);\n"
                "}\n"
                f"hipError_t eri_launchc_{tag}_{prec}(const EriArgs& a, hipStream_t s) {{\n"
                f"  return eri_launch_coop<{args}>(a, s);\n"
                "}\n}  // namespace qcmd\n"
            )
            if not os.path.exists(path) or open(path).read() != text:
                with open(path, "w") as fh:
                    fh.write(text)
            # rough cost for scheduling: total angular momentum, fp64 first
            out.append((sum(c) * 2 + (prec == "f64"), path))
        ch = GRAD_PRIM_CHUNK.get(tag, (1, 1))
        ych = GRAD_YPRIM_CHUNK.get(tag, (1, 1))
        unsupported = "  (void)a;\n  (void)s;\n  return hipErrorNotSupported;\n"

        def fn(name, body):
            return (f"hipError_t grad2e_{name}_{tag}(const Grad2eArgs& a, hipStream_t s) {{\n"
                    + body + "}\n")

        args = f"{c[0]}, {c[1]}, {c[2]}, {c[3]}"
        ncd = (c[2] + 1) * (c[2] + 2) // 2 * ((c[3] + 1) * (c[3] + 2) // 2)
        part_units = {}

        def yprim_entry(name, real, chunk, built):
            """The G-first launcher of one precision: its chunks spread over
            parts of at most YPRIM_CHUNKS_PER_UNIT instantiations, each part
            its own translation unit, called in order by the entry point."""
            if not built:
                return fn(name, unsupported)
            nch = -(-ncd // chunk)
            per = YPRIM_CHUNKS_PER_UNIT
            if nch <= per:
                return fn(name, f"  return grad2e_launch_yprim<{args}, {real}, {chunk}>(a, s);\n")
            body, decls = "", ""
            for k, c0 in enumerate(range(0, nch, per)):
                cda, cdb = c0 * chunk, min(ncd, (c0 + per) * chunk)
                pname = f"{name}_{tag}_part{k}"
                decls += f"hipError_t grad2e_{pname}(const Grad2eArgs& a, hipStream_t s);\n"
                part_units[f"_y{real[0]}{k}"] = (
                    f"hipError_t grad2e_{pname}(const Grad2eArgs& a, hipStream_t s) {{\n"
                    f"  return grad2e_launch_yprim_range<{args}, {real}, {chunk}, {cda}, {cdb}>(a, s);\n}}\n")
                body += f"  if (hipError_t e = grad2e_{pname}(a, s); e != hipSuccess) return e;\n"
            return decls + fn(name, body + "  return hipSuccess;\n")

        # one unit per kernel form, so that the heavy instantiations of one
        # class compile in parallel
        units = {
            "": fn("launchw", f"  return grad2e_launch_wave<{args}>(a, s);\n")
                + fn("launchw32", f"  return grad2e_launch_wave<{args}, float>(a, s);\n")
                + fn("launcht", f"  return grad2e_launch_thread<{args}>(a, s);\n"
                     if tag in GRAD_PER_THREAD else unsupported)
                + fn("launcht32", f"  return grad2e_launch_thread<{args}, float>(a, s);\n"
                     if tag in GRAD_PER_THREAD_F32 else unsupported),
            "_p": fn("launchp", f"  return grad2e_launch_prim<{args}, double, {ch[0]}>(a, s);\n"
                     if tag in GRAD_PRIM else unsupported)
                + fn("launchp32", f"  return grad2e_launch_prim<{args}, float, {ch[1]}>(a, s);\n"
                     if tag in GRAD_PRIM_F32 else unsupported),
            "_y": yprim_entry("launchy", "double", ych[0], tag in GRAD_YPRIM)
                + yprim_entry("launchy32", "float", ych[1], tag in GRAD_YPRIM_F32),
            "_m": fn("launchwm", f"  return grad2e_launch_wave<{args}, double, true>(a, s);\n"
                     if tag in GRAD_WAVE_MFMA else unsupported)
                + fn("launchwm32", f"  return grad2e_launch_wave<{args}, float, true>(a, s);\n"
                     if tag in GRAD_WAVE_MFMA else unsupported),
        }
        units.update(part_units)
        for suffix, body in units.items():
            path = os.path.join(GEN, f"grad2e_{tag}{suffix}.hip")
            text = '#include "grad_prim.h"\nnamespace qcmd {\n' + body + "}  // namespace qcmd\n"
            if not os.path.exists(path) or open(path).read() != text:
                with open(path, "w") as fh:
                    fh.write(text)
            heavy = ((suffix == "_p" and (tag in GRAD_PRIM or tag in GRAD_PRIM_F32))
                     or (suffix.startswith("_y") and tag in GRAD_YPRIM) or suffix == "")
            out.append((sum(c) * 2 + (2 if heavy else 0), path))
    return out


def _headers_digest():
    h = hashlib.sha256()
    for name in sorted(os.listdir(HERE)):
        if name.endswith(".h"):
            h.update(open(os.path.join(HERE, name), "rb").read())
    h.update(open(os.path.join(REPO, "include", "qcmd.h"), "rb").read())
    h.update(" ".join(FLAGS).encode())
    return h.hexdigest()


def _compile(src, reuse, hdr):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    digest = hashlib.sha256(open(src, "rb").read() + hdr.encode()).hexdigest()
    stamp = obj + ".sha"
    if reuse and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == digest:
        return obj, 0.0, True
    t = time.time()
    r = subprocess.run([HIPCC, *FLAGS, "-c", src, "-o", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    with open(stamp, "w") as fh:
        fh.write(digest)
    return obj, time.time() - t, False


def build(jobs: int | None = None, reuse: bool = False, verbose: bool = True) -> str:
    t0 = time.time()
    gen = sorted(_generate(), reverse=True)
    srcs = [p for _, p in gen] + [os.path.join(HERE, s) for s in STATIC_SOURCES
                                  if os.path.exists(os.path.join(HERE, s))]
    hdr = _headers_digest()
    jobs = jobs or max(1, os.cpu_count() or 1)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, reuse, hdr): s for s in srcs}
        for f in cf.as_completed(futs):
            obj, dt, cached = f.result()
            objs.append(obj)
            if verbose and not cached:
                print(f"  [{time.time() - t0:6.1f}s] {os.path.basename(obj)} {dt:.1f}s", flush=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *sorted(objs)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--reuse", action="store_true")
    a = ap.parse_args()
    try:
        build(a.jobs, a.reuse)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
