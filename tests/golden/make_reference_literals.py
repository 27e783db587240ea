#!/usr/bin/env python3
"""Write tests/golden/reference_literals.json: the numeric literals of the
reference's hot-path functions, for the oracle's constant pins.

For each named region of the reference sources (file and line ranges, given
below with what they hold) the fixture records the numeric literals in the
order they appear, with comments removed, and for the files' `#define NAME
<numeric expression>` lines the evaluated value.  Only numbers are written:
the fixture holds no source text.  tests/test_oracle_constants.py compares the
literals of the oracle's restatement of each region (oracle/ryr_oracle.c) with
these, in order, so a mistyped, missing or swapped constant in the oracle
fails.

Run in a checkout that has the reference next to it (it reads the sources
as text; nothing is compiled or executed):
    python tests/golden/make_reference_literals.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# region name -> [(file, first line, last line), ...]  (1-based, inclusive)
REGIONS = {
    # subroutine_cuda.c: NCX with its allosteric factor
    "ncx": [("subroutine_cuda.c", 47, 75)],
    # SERCA uptake and NSR leak
    "uptake": [("subroutine_cuda.c", 128, 137)],
    "leak": [("subroutine_cuda.c", 139, 146)],
    # CSQN rapid-buffer factor and its steady-state pool
    "luminal": [("subroutine_cuda.c", 161, 180)],
    "luminal_ss": [("subroutine_cuda.c", 182, 199)],
    # SR release flux and the GHK single-channel current
    "release": [("subroutine_cuda.c", 222, 228)],
    "lcccurrent": [("subroutine_cuda.c", 231, 247)],
    # 16-state LCC Markov chain, default and iso2 gate parameters
    "lcc_markov": [("reesica.h", 14, 116)],
    # whole-cell scaling of the sums, then the membrane currents, gates,
    # [Na]i and dV/dt (stimulus handled apart)
    "membrane": [("cuda_Cell_Detailed.cu", 350, 351), ("cuda_Cell_Detailed.cu", 452, 452),
                 ("cuda_Cell_Detailed.cu", 456, 456), ("cuda_Cell_Detailed.cu", 458, 676),
                 ("cuda_Cell_Detailed.cu", 700, 731)],
    # host initial conditions of the membrane state
    "membrane_ics": [("cuda_Cell_Detailed.cu", 353, 365)],
}
# `#define` blocks, scanned in this order (the buffer file's volumes use the
# main file's Vi and Vs), under the switches of the published build: run.sh:5
# builds `make il`, whose recipe (Makefile:20-22) passes -Diso2 -Dlqt; the
# file's own active switches (lumgate, outputlinescan) are added as they are
# met.  The values are those of that build (e.g. svikr 0 under lqt).
DEFINE_FILES = [("cuda_Cell_Detailed.cu", 29, 146), ("subroutine_cuda.c", 1, 45), ("buffers.c", 1, 53)]
BUILD_SWITCHES = {"iso2", "lqt"}

NUM = re.compile(r"(?<![\w.])(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?(?![\w.])")


def strip_comments(text):
    """Remove /* */ and // comments, keeping line breaks so line numbers hold."""
    out, i, n = [], 0, len(text)
    while i < n:
        if text.startswith("/*", i):
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
            out.append("\n" * text.count("\n", i, j))
            i = j
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        else:
            out.append(text[i])
            i += 1
    return "".join(out)


def literals(lines):
    found = []
    for ln in lines:
        if ln.lstrip().startswith("#"):        # preprocessor lines carry no arithmetic here
            continue
        found.extend(NUM.findall(ln))
    return found


def eval_numeric(expr, known):
    e = expr.strip()
    for name in sorted(known, key=len, reverse=True):
        e = re.sub(rf"\b{name}\b", repr(known[name]), e)
    if not re.fullmatch(r"[0-9eE.+\-*/() ]+", e):
        return None
    try:
        return float(eval(e, {"__builtins__": {}}))
    except Exception:
        return None


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = {}
    for f in {r[0] for v in REGIONS.values() for r in v} | {d[0] for d in DEFINE_FILES}:
        with open(os.path.join(ref, f)) as fh:
            src[f] = strip_comments(fh.read()).split("\n")
    out = {"regions": {}, "defines": {}}
    for name, parts in REGIONS.items():
        lits = []
        for f, a, b in parts:
            lits.extend(literals(src[f][a - 1:b]))
        out["regions"][name] = {"source": [f"{f}:{a}-{b}" for f, a, b in parts], "literals": lits}
    known, flags = {}, set(BUILD_SWITCHES)
    for f, a, b in DEFINE_FILES:
        active = [True]
        for k, ln in enumerate(src[f][a - 1:b], start=a):
            t = ln.strip()
            m = re.match(r"#\s*(ifdef|ifndef)\s+(\w+)", t)
            if m:
                on = (m.group(2) in flags) == (m.group(1) == "ifdef")
                active.append(active[-1] and on)
                continue
            if re.match(r"#\s*else\b", t):
                outer = active[-2] if len(active) > 1 else True
                active[-1] = outer and not active[-1]
                continue
            if re.match(r"#\s*endif\b", t):
                active.pop()
                continue
            if not active[-1]:
                continue
            m = re.match(r"#\s*define\s+(\w+)\s*(.*?)\s*$", t)
            if not m:
                continue
            if not m.group(2):
                flags.add(m.group(1))
                continue
            val = eval_numeric(m.group(2), known)
            if val is not None:
                known[m.group(1)] = val
                out["defines"][m.group(1)] = {"value": val, "at": f"{f}:{k}"}
    out["switches"] = sorted(flags)
    path = os.path.join(HERE, "reference_literals.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}: {len(out['regions'])} regions, {len(out['defines'])} defines")


if __name__ == "__main__":
    main()
