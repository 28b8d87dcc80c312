"""Builds libasdfft.so (HIP kernels for gfx950 + C++ host library) in-tree.

    python -m sip_amd.build [--force] [--jobs N]

Kernel instances compile in parallel (one translation unit per (variant, EMAX)
entry of ASDFFT_FOR_EACH_INSTANCE); host files compile with g++ against the
HIP runtime headers; everything links with hipcc into sip_amd/lib/libasdfft.so.
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libasdfft.so")
TORCH_LIB = os.path.join(LIBDIR, "libasdfft_torch.so")
ARCH = os.environ.get("ASDFFT_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # packed FP32 instructions are kept out of the generated code on purpose
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
              f"-I{ROCM}/include", f"-I{INCLUDE}", f"-I{CSRC}"]
HOST_SOURCES = ["acl_tensor.cpp", "twiddles.cpp", "planner.cpp", "ragged.cpp", "server.cpp", "serve_api.cpp",
                "fft_api.cpp"]
KERNEL_HEADERS = ["stockham_impl.h", "fft_kernel_args.h", "kernels.h", "radix_consts.h", "ragged.h", "mfma_fft.h", "shift.h"]
# hand-written kernel translation units besides the generated instances
KERNEL_SOURCES = ["dft_direct.hip", "mfma_fft.hip", "shift_kernel.hip", "istft_kernel.hip", "real_aux.hip",
                  "static_scratch.hip", "wave_fft.hip"]
# the ragged kernel builds once per register budget (one translation unit each)
RAGGED_EMAX = [16, 32, 48]


def instances():
    text = open(os.path.join(CSRC, "kernels.h")).read()
    body = text[text.index("#define ASDFFT_FOR_EACH_INSTANCE"):]
    return [(int(v), int(e)) for v, e in re.findall(r"X\((\d+),\s*(\d+)\)", body)]


def static_instances():
    """Rows of static_instances.txt: (variant, emax, n, threads, tpb, radices)."""
    out = []
    for line in open(os.path.join(CSRC, "static_instances.txt")):
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        v = [int(x) for x in line.split()]
        out.append((v[0], v[1], v[2], v[3], v[4], tuple(v[5:])))
    return out


def generate_static_sources():
    """One .hip per specialised kernel plus the host registry table (into _build/gen)."""
    gen = os.path.join(BUILD, "gen")
    os.makedirs(gen, exist_ok=True)
    entries = static_instances()
    hips = []
    for k, (v, e, n, t, tpb, rad) in enumerate(entries):
        src = os.path.join(gen, f"static_{k}.hip")
        body = (
            "// generated from static_instances.txt\n"
            '#include "stockham_impl.h"\n'
            "namespace asdfft {\n"
            f"hipError_t launch_static_entry_{k}(const LaunchShape& s, const FftArgs& a, hipStream_t st)\n"
            "{\n"
            f"    return launch_static<{v}, {e}, {n}, {t}, {max(tpb, 1)}, {', '.join(map(str, rad))}>(s, a, st);\n"
            "}\n"
            "}  // namespace asdfft\n"
        )
        if not os.path.exists(src) or open(src).read() != body:
            with open(src, "w") as f:
                f.write(body)
        hips.append(src)
    reg = ["// generated from static_instances.txt", '#include "kernels.h"', "namespace asdfft {"]
    for k in range(len(entries)):
        reg.append(f"hipError_t launch_static_entry_{k}(const LaunchShape&, const FftArgs&, hipStream_t);")
    reg.append("const StaticEntry kStaticInstances[] = {")
    for k, (v, e, n, t, tpb, rad) in enumerate(entries):
        reg.append(f"    {{{v}, {e}, {n}, {t}, {tpb}, {len(rad)}, {{{', '.join(map(str, rad))}}}, "
                   f"&launch_static_entry_{k}}},")
    if not entries:
        reg.append("    {0, 0, 0, 0, 0, 0, {0}, nullptr},")
    reg.append("};")
    reg.append(f"const int kNumStaticInstances = {len(entries)};")
    reg.append("}  // namespace asdfft")
    reg_src = os.path.join(gen, "static_registry.cpp")
    text = "\n".join(reg) + "\n"
    if not os.path.exists(reg_src) or open(reg_src).read() != text:
        with open(reg_src, "w") as f:
            f.write(text)
    return hips, reg_src


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_torch_ops(force: bool = False) -> str:
    """libasdfft_torch.so: the torch_sip operators (torch_ops.cpp) over libasdfft.so."""
    import torch
    tdir = os.path.dirname(torch.__file__)
    src = os.path.join(CSRC, "torch_ops.cpp")
    deps = [src, LIB] + [os.path.join(INCLUDE, h) for h in os.listdir(INCLUDE) if h.endswith(".h")]
    if not force and not _newer(TORCH_LIB, deps):
        return TORCH_LIB
    abi = int(torch.compiled_with_cxx11_abi())
    tmp = TORCH_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DUSE_ROCM",
          "-D__HIP_PLATFORM_AMD__", f"-I{tdir}/include", f"-I{tdir}/include/torch/csrc/api/include",
          f"-I{ROCM}/include", f"-I{INCLUDE}", src, "-o", tmp, f"-L{tdir}/lib", "-lc10", "-lc10_hip", "-ltorch",
          "-ltorch_cpu", "-ltorch_hip", f"-L{LIBDIR}", "-lasdfft", f"-L{ROCM}/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined"])
    os.replace(tmp, TORCH_LIB)
    return TORCH_LIB


def host_driver(driver: str, out: str, sanitize: str | None = None, jobs: int | None = None) -> str:
    """Links a host program (driver) against the library's host sources compiled for it,
    optionally instrumented (sanitize: "address,undefined" or "thread"), and the device
    objects of the last build().  The host sources are recompiled for each instrumentation
    into their own folder under _build; nothing of libasdfft.so changes."""
    dev_objs = [os.path.join(BUILD, f"stockham_v{v}_e{e}.o") for v, e in instances()]
    static_hips, reg_src = generate_static_sources()
    dev_objs += [src[:-4] + ".o" for src in static_hips]
    dev_objs += [os.path.join(BUILD, f"ragged_kernel_e{e}.o") for e in RAGGED_EMAX]
    dev_objs += [os.path.join(BUILD, f.replace(".hip", ".o")) for f in KERNEL_SOURCES]
    if not all(os.path.exists(o) for o in dev_objs):
        build(jobs=jobs)
    tag = (sanitize or "plain").replace(",", "_")
    odir = os.path.join(BUILD, f"host_{tag}")
    os.makedirs(odir, exist_ok=True)
    flags = ["-O1", "-g", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{INCLUDE}",
             f"-I{CSRC}", "-fno-omit-frame-pointer"]
    if sanitize:
        flags += [f"-fsanitize={sanitize}", "-fno-sanitize-recover=all"]
    host_hdr = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    host_hdr += [os.path.join(INCLUDE, h) for h in os.listdir(INCLUDE) if h.endswith(".h")]
    tasks, objs = [], []
    for src in [os.path.join(CSRC, f) for f in HOST_SOURCES] + [reg_src, driver]:
        obj = os.path.join(odir, os.path.basename(src).rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        if _newer(obj, [src] + host_hdr):
            tasks.append(["g++"] + flags + ["-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs or min(16, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, tasks))
    if tasks or _newer(out, objs + dev_objs):
        link = ["g++", "-o", out] + objs + dev_objs + [f"-L{ROCM}/lib", "-lamdhip64", "-lpthread",
                                                       f"-Wl,-rpath,{ROCM}/lib"]
        if sanitize:
            link.insert(1, f"-fsanitize={sanitize}")
        _run(link)
    return out


def build(force: bool = False, jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_radix_consts.py")], check=True)
    hdr = [os.path.join(CSRC, h) for h in KERNEL_HEADERS]
    jobs = jobs or min(16, os.cpu_count() or 4)
    tasks = []
    objs = []
    for v, e in instances():
        obj = os.path.join(BUILD, f"stockham_v{v}_e{e}.o")
        src = os.path.join(CSRC, "stockham_inst.hip")
        objs.append(obj)
        if force or _newer(obj, [src] + hdr):
            tasks.append(["hipcc"] + HIPCC_FLAGS + [f"-DASDFFT_VARIANT={v}", f"-DASDFFT_EMAX={e}", "-c", src,
                                                    "-o", obj])
    static_hips, reg_src = generate_static_sources()
    for src in static_hips:
        obj = src[:-4] + ".o"
        objs.append(obj)
        if force or _newer(obj, [src] + hdr):
            tasks.append(["hipcc"] + HIPCC_FLAGS + ["-c", src, "-o", obj])
    reg_obj = reg_src[:-4] + ".o"
    objs.append(reg_obj)
    if force or _newer(reg_obj, [reg_src] + hdr):
        tasks.append(["g++"] + HOST_FLAGS + ["-c", reg_src, "-o", reg_obj])
    for e in RAGGED_EMAX:
        obj = os.path.join(BUILD, f"ragged_kernel_e{e}.o")
        src = os.path.join(CSRC, "ragged_kernel.hip")
        objs.append(obj)
        if force or _newer(obj, [src] + hdr):
            tasks.append(["hipcc"] + HIPCC_FLAGS + [f"-DASDFFT_RAGGED_EMAX={e}", "-c", src, "-o", obj])
    for f in KERNEL_SOURCES:
        obj = os.path.join(BUILD, f.replace(".hip", ".o"))
        objs.append(obj)
        if force or _newer(obj, [os.path.join(CSRC, f)] + hdr):
            tasks.append(["hipcc"] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, f), "-o", obj])
    host_hdr = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    host_hdr += [os.path.join(INCLUDE, h) for h in os.listdir(INCLUDE) if h.endswith(".h")]
    for f in HOST_SOURCES:
        obj = os.path.join(BUILD, f.replace(".cpp", ".o"))
        objs.append(obj)
        if force or _newer(obj, [os.path.join(CSRC, f)] + host_hdr):
            tasks.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, f), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, tasks):
            if verbose and out.strip():
                print(out)
    if force or tasks or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
             [f"-L{ROCM}/lib", "-lamdhip64", "-Wl,--no-undefined"])
        os.replace(tmp, LIB)
    build_torch_ops(force)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.jobs, verbose=a.verbose))


if __name__ == "__main__":
    main()
