"""Build the MI355X P3M library in-tree with hipcc (gfx950).

Produces ``espresso_amd/lib/libespresso_amd_p3m.so`` from the sources in
``espresso_amd/csrc``.  The library exports the C ABI declared in
``include/espresso_amd_p3m.h``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libespresso_amd_p3m.so"
LIB_PATH = os.path.join(LIBDIR, LIBNAME)

SOURCES = ["p3m_host.cpp", "p3m_tuning.cpp", "p3m_capi.cpp", "p3m_solver.hip",
           "p3m_batch.hip", "p3m_serve.cpp", "p3m_short_range.hip", "p3m_elc.hip", "p3m_dipolar.hip"]

# -packed-fp32-ops keeps the compiler from emitting packed FP32 VALU forms
# (v_pk_*_f32); hipcc's host pass warns that it does not know the feature.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++20",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-warning-option",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# host-code sanitizers (the device code is built as usual): every
# -fsanitize goes to the host pass only
SANITIZE_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
SANITIZE_LIB = os.path.join(LIBDIR, "sanitize", LIBNAME)


def build(verbose=False, jobs=None, sanitize=False):
    """Compile every source (always) and link the shared library.  With
    sanitize, the host code is instrumented with AddressSanitizer and
    UndefinedBehaviorSanitizer and the library goes to lib/sanitize/."""
    target = SANITIZE_LIB if sanitize else LIB_PATH
    os.makedirs(os.path.dirname(target), exist_ok=True)
    objdir = os.path.join(HERE, "_build_sanitize" if sanitize else "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    procs = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmd = [cc] + lang + HIPFLAGS + (SANITIZE_FLAGS if sanitize else []) + inc + [
            "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write(f"--- {src} ---\n{text}\n")
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))
    tmp = target + ".tmp"
    link_san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host",
                "-fsanitize=undefined", "-shared-libsan"] if sanitize else []
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC"] + link_san + ["-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, sanitize="--sanitize" in sys.argv))
