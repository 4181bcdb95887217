"""CPU build of libkcore_hip for tests: the library's host sources (kcore_hip.hip, serve.hip,
tune.cpp) compiled unchanged against the stand-in HIP runtime of tests/hostemu/hip, with every
kernel run by the wave64 emulator of tests/emu. Only the `kernel<<<grid, block, shmem,
stream>>>(args)` launches are rewritten, into a call that enqueues the emulated launch.

    python tests/hostemu/cpu_lib.py <out.so>     builds the shared library

Used by tests/test_hostemu.py and by the opt-in pytest plugin cpu_lib_plugin.py. Nothing built
here is a GPU result."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
CLANG = "/opt/rocm/llvm/bin/clang++"
CC = CLANG if os.path.exists(CLANG) else "clang++"
HIP_SOURCES = ("kcore_hip.hip", "serve.hip", "tune.cpp")


def _split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    parts.append(cur.strip())
    return parts


def rewrite_launches(text):
    """kernel<...><<<grid, block, shmem, stream>>>(args);  ->  kcemu_launch(grid, block, stream, kernel<...>, args);
    Returns the new text and the number of launches rewritten."""
    out, i, n = [], 0, 0
    while True:
        k = text.find("<<<", i)
        if k < 0:
            out.append(text[i:])
            return "".join(out), n
        j = k
        if text[j - 1] == ">":  # template arguments of the kernel
            depth, j = 0, j - 1
            while True:
                depth += {">": 1, "<": -1}.get(text[j], 0)
                if depth == 0:
                    break
                j -= 1
        m = j
        while m > 0 and (text[m - 1].isalnum() or text[m - 1] in "_:"):
            m -= 1
        e = text.index(">>>", k)
        cfg = _split_top(text[k + 3:e])
        assert len(cfg) == 4 and text[e + 3] == "(", text[m:e + 10]
        depth, p = 0, e + 3
        while True:
            depth += {"(": 1, ")": -1}.get(text[p], 0)
            if depth == 0:
                break
            p += 1
        args = text[e + 4:p].strip()
        out.append(text[i:m])
        out.append(f"kcemu_launch({cfg[0]}, {cfg[1]}, {cfg[3]}, {text[m:k]}" + (f", {args})" if args else ")"))
        i = p + 1
        n += 1


def build_objects(work, extra=()):
    """Copies the sources into `work`, rewrites the launches and compiles them (position
    independent; `extra`: more compiler flags, e.g. a sanitizer). Returns the object files."""
    csrc = os.path.join(work, "kcoregpu_amd", "csrc")
    shutil.copytree(os.path.join(ROOT, "kcoregpu_amd", "csrc"), csrc, ignore=shutil.ignore_patterns("*.so", "*.o"))
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(work, "include"))
    objs = []
    for f in HIP_SOURCES:
        p = os.path.join(csrc, f)
        with open(p) as fh:
            text, _ = rewrite_launches(fh.read())
        assert "<<<" not in text
        with open(p, "w") as fh:
            fh.write(text)
        o = p + ".o"
        subprocess.run([CC, "-O1", "-std=c++17", "-fPIC", "-pthread", *extra, "-I", HERE, "-x", "c++", "-c", p,
                        "-o", o], check=True)
        objs.append(o)
    return objs


def build_shared(objs, out):
    subprocess.run([CC, "-shared", "-pthread", "-Wl,-Bsymbolic", *objs, "-o", out], check=True)
    return out


def main(argv):
    if len(argv) != 2:
        print(__doc__)
        return 2
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        build_shared(build_objects(td), os.path.abspath(argv[1]))
    print(argv[1])
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
