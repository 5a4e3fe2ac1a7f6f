"""Per-kernel resource report of libmatfun_hip.so's sources for gfx950, from the compiler's own accounting
(-Rpass-analysis=kernel-resource-usage): VGPRs, AGPRs, SGPRs, scratch bytes per lane, VGPR / SGPR spills, LDS bytes
per workgroup and the occupancy in wavefronts per SIMD that those allow.

    python -m matfun_amd.resources > report.md     (compiles every source once more, device code only, with the
                                                    library's flags; a few minutes: the batch kernels are large)

The compile writes its objects to a temporary directory; the library itself is not touched.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

from matfun_amd import build as B

FIELDS = [("TotalSGPRs", "SGPRs"), ("VGPRs", "VGPRs"), ("AGPRs", "AGPRs"), ("ScratchSize [bytes/lane]", "scratch B/lane"),
          ("Occupancy [waves/SIMD]", "waves/SIMD"), ("SGPRs Spill", "SGPR spill"), ("VGPRs Spill", "VGPR spill"),
          ("LDS Size [bytes/block]", "LDS B/WG")]
_REMARK = re.compile(r"remark:\s+(Function Name|[A-Za-z ]+(?:\[[^\]]*\])?):\s*(\S+) \[-Rpass-analysis")


def demangle(names):
    tool = shutil.which("llvm-cxxfilt") or shutil.which("c++filt")
    if not tool:
        return list(names)
    r = subprocess.run([tool], input="\n".join(names), capture_output=True, text=True)
    out = r.stdout.splitlines()
    return out if len(out) == len(names) else list(names)


def parse(text):
    """[(mangled name, {field: value})] in the order of the remarks"""
    kernels = []
    for m in _REMARK.finditer(text):
        key, val = m.group(1).strip(), m.group(2)
        if key == "Function Name":
            kernels.append((val, {}))
        elif kernels:
            kernels[-1][1][key] = val
    return kernels


def compile_source(src, tmp):
    cmd = [B.hipcc(), *B.flags(["-Rpass-analysis=kernel-resource-usage", "--offload-device-only"] +
                               B.PER_FILE.get(src, [])),
           "-c", os.path.join(B.CSRC, src), "-o", os.path.join(tmp, os.path.splitext(src)[0] + ".o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return src, parse(r.stderr)


def _split_args(name):
    """('base', [template arguments]) of a kernel name (top-level commas only), or (name, None)"""
    if not name.endswith(">") or "<" not in name:
        return name, None
    i = name.index("<")
    args, depth, cur = [], 0, ""
    for ch in name[i + 1:-1]:
        if ch == "," and depth == 0:
            args.append(cur.strip())
            cur = ""
            continue
        depth += ch == "<"
        depth -= ch == ">"
        cur += ch
    args.append(cur.strip())
    return name[:i], args


def _collapse(rows):
    """rows (source, kernel, values) merged where the values agree and the kernels differ in one template argument:
    smf_kernel<{0,1,2,3}, true, ...> (repeated until nothing merges)"""
    items = []
    for src, name, vals in rows:
        base, args = _split_args(name)
        items.append((src, base, None if args is None else [frozenset([a]) for a in args], vals))
    merged = True
    while merged:
        merged = False
        for i in range(len(items)):
            for j in range(i + 1, len(items)):
                a, b = items[i], items[j]
                if a[0] != b[0] or a[1] != b[1] or a[3] != b[3] or a[2] is None or b[2] is None or len(a[2]) != len(b[2]):
                    continue
                diff = [k for k in range(len(a[2])) if a[2][k] != b[2][k]]
                if len(diff) == 1:
                    k = diff[0]
                    args = list(a[2])
                    args[k] = a[2][k] | b[2][k]
                    items[i] = (a[0], a[1], args, a[3])
                    del items[j]
                    merged = True
                    break
            if merged:
                break

    def fmt(vals):
        v = sorted(vals, key=lambda x: (len(x), x))
        return v[0] if len(v) == 1 else "{" + ",".join(v) + "}"
    return [(src, base if args is None else base + "<" + ", ".join(fmt(a) for a in args) + ">", vals)
            for src, base, args, vals in items]


def report(sources=None):
    sources = sources or B.SOURCES
    with tempfile.TemporaryDirectory() as tmp, ThreadPoolExecutor(min(len(sources), os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda s: compile_source(s, tmp), sources))
    rows, spills = [], []
    for src, ks in results:
        names = demangle([k for k, _ in ks])
        for name, (_, f) in zip(names, ks):
            short = name.replace("void ", "").replace("(anonymous namespace)::", "").replace("mf::", "")
            short = re.sub(r"\(.*\)$", "", short)
            vals = tuple(f.get(k, "?") for k, _ in FIELDS)
            rows.append((src, short, vals))
            if f.get("ScratchSize [bytes/lane]", "0") != "0" or f.get("VGPRs Spill", "0") != "0":
                spills.append((src, short, f.get("ScratchSize [bytes/lane]"), f.get("VGPRs Spill")))
    lines = ["| source | kernels | VGPRs | AGPRs | waves/SIMD | LDS B/WG | with scratch |", "|---|---|---|---|---|---|---|"]
    for src in sources:
        sel = [v for s_, _, v in rows if s_ == src]
        if not sel:
            continue
        col = lambda i: [int(v[i]) for v in sel if v[i].isdigit()]
        rng = lambda xs: f"{min(xs)}-{max(xs)}" if xs and min(xs) != max(xs) else (str(xs[0]) if xs else "?")
        lines.append(f"| {src} | {len(sel)} | {rng(col(1))} | {rng(col(2))} | {rng(col(4))} | {rng(col(7))} | "
                     f"{sum(1 for v in sel if v[3] != '0')} |")
    lines += ["", "| source | kernel | " + " | ".join(h for _, h in FIELDS) + " |",
              "|---|---|" + "---|" * len(FIELDS)]
    for src, name, vals in sorted(_collapse(rows)):
        lines.append(f"| {src} | `{name}` | " + " | ".join(vals) + " |")
    lines += ["", f"kernels: {len(rows)}; with scratch or VGPR spills: {len(spills)}"]
    lines += [f"- {s} `{k}`: scratch {sc} B/lane, {v} VGPRs spilled" for s, k, sc, v in sorted(spills)]
    return "\n".join(lines)


if __name__ == "__main__":
    print(report(sys.argv[1:] or None))
