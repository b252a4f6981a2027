"""Undefined-behaviour and signed-wrap checks of the CPU paths (test
infrastructure: run by tests/test_ub_builds.py, never by the product).

`make -C oracle checks` builds the oracle and the reference's sources (with
the shim) twice more: under UBSan (-fsanitize=undefined, -O1 -g) and with
-fwrapv at -O3.  This module runs one fixed set of workloads -- the golden
fixtures' inputs plus random full-range int16 blocks -- through one build in
a child process and saves every output:

* under UBSan, the child's log lists each source line whose operation was
  undefined on these inputs (signed overflow, shifts of negative values,
  out-of-range conversions);
* the -fwrapv build's outputs, compared with the -O3 build's, show where the
  optimiser's use of "overflow cannot happen" changes a result (the golden
  fixtures hold the -O3 results, which the oracle and the kernels follow).

    python -m oracle.ub_check run <build> <impl> <out.npz>   (child)
    python -m oracle.ub_check report                         (summary)
"""
from __future__ import annotations

import collections
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(ROOT, "tests", "golden")
BUILDS = {
    "O3": {"oracle": os.path.join(HERE, "liboracle.so"),
           "ref": os.path.join(HERE, "_ref", "libevxref.so")},
    "ubsan": {"oracle": os.path.join(HERE, "_ref", "ubsan", "liboracle.so"),
              "ref": os.path.join(HERE, "_ref", "ubsan", "libevxref.so")},
    "fwrapv": {"oracle": os.path.join(HERE, "_ref", "fwrapv", "liboracle.so"),
               "ref": os.path.join(HERE, "_ref", "fwrapv", "libevxref.so")},
}


def available():
    return all(os.path.exists(p) for b in BUILDS.values() for p in b.values())


def _load(name):
    return np.load(os.path.join(GOLD, name))


def _planes(g, tag):
    return tuple(g[f"{tag}_{c}"] for c in "yuv")


def workloads(impl, is_ref):
    """Yield (name, array) outputs of the fixed workload set."""
    from oracle.oracle import REF_FIELDS, Picture, SAO_DTYPE

    m = _load("metrics.npz")
    yield "metrics", np.array([impl.metrics(a, b) for a, b in zip(m["a"], m["b"])], np.int64)
    yield "hadamard", np.array([impl.hadamard(h.reshape(4, 4)) for h in m["had_in"]])

    d = _load("dctquant.npz")
    rng = np.random.default_rng(77)
    src = np.concatenate([d["src"], rng.integers(-32768, 32768, (200, 384))]).astype(np.int16)
    pred = np.concatenate([d["pred"], rng.integers(-32768, 32768, (200, 384))]).astype(np.int16)
    co = np.array([impl.sub_transform(a, b) for a, b in zip(src, pred)])
    yield "sub_transform", co
    yield "variance2", np.array([impl.variance2(c) for c in co], np.int64)
    yield "block_qp", np.array([[impl.block_qp(q, c) for q in (0, 8, 16, 31, 63)] for c in co])
    yield "quantize", np.array([[impl.quantize(qp, c) for qp in (1, 8, 17, 40, 63)] for c in co])
    yield "metrics_wide", np.array([impl.metrics(a, b) for a, b in zip(src, pred)], np.int64)

    r = _load("recon.npz")
    lv = np.concatenate([r["dq_levels"], rng.integers(-32768, 32768, (100, 384))]).astype(np.int16)
    qp = np.concatenate([r["dq_qp"], rng.integers(1, 64, 100)])
    yield "dequantize", np.array([impl.dequantize(int(q), l) for q, l in zip(qp, lv)])
    cw = np.concatenate([r["it_coeffs"], rng.integers(-32768, 32768, (100, 384))]).astype(np.int16)
    pw = np.concatenate([r["it_pred"], rng.integers(-32768, 32768, (100, 384))]).astype(np.int16)
    yield "idct_add", np.array([impl.idct_add(c, p) for c, p in zip(cw, pw)])

    g = _load("interp.npz")
    for tag in ("u8", "wide"):
        P = Picture(*_planes(g, tag))
        p = g[f"{tag}_params"]
        if is_ref:
            out = impl.interp_many(P, p[:, 0], p[:, 1], p[:, 2], p[:, 3], p[:, 4])
        else:
            out = np.array([impl.interp(P, *map(int, row)) for row in p])
        yield f"interp_{tag}", out

    s = _load("search.npz")
    for tag in ("s64", "s128", "s128c"):
        cur, ref = Picture(*_planes(s, f"{tag}_cur")), Picture(*_planes(s, f"{tag}_ref"))
        rows = []
        for x, y, q, pmx, pmy in s[f"{tag}_params"]:
            res = impl.inter_search(cur.block(int(x), int(y)), ref, int(x), int(y), int(q),
                                    int(pmx), int(pmy))
            if is_ref:
                rows.append([int(res[k]) for k in ("block_type", "motion_x", "motion_y",
                                                   "sp_pred", "sp_amount", "sp_index", "sad")])
            else:
                rows.append([res.block_type, res.motion_x, res.motion_y, res.sp_enabled,
                             res.sp_amount, res.sp_index, res.sad])
        yield f"search_{tag}", np.array(rows, np.int64)

    f = _load("frames.npz")
    pics = [Picture(*_planes(f, f"f{k}")) for k in range(6)]
    prev = None
    for tag, ftype, q, cur, refs in (("p1", 1, 8, 1, [0]), ("p2", 1, 8, 2, [1, 0]),
                                     ("p4", 1, 12, 5, [4, 3, 2, 1]), ("b", 2, 9, 3, [2, 4])):
        rec, co_, lv_ = impl.analyze_frame(ftype, q, pics[cur], [pics[k] for k in refs],
                                           prev=prev if ftype == 1 else None)
        if ftype == 1:
            prev = rec
        yield f"frame_{tag}_records", np.stack([rec[n].astype(np.int64) for n in REF_FIELDS])
        if not is_ref:
            yield f"frame_{tag}_diagnostics", np.stack([rec[n].astype(np.int64) for n in (
                "mad", "cost_bits", "mad_b", "cost_b_bits", "n_int", "n_axial", "n_diag")])
        yield f"frame_{tag}_coeffs", co_
        yield f"frame_{tag}_levels", lv_
        recon = impl.reconstruct_frame(ftype, q, pics[cur], [pics[k] for k in refs], rec,
                                       r[f"{tag}_levels_perturbed"])
        yield f"recon_{tag}", np.concatenate([a.ravel() for a in recon])
        db = impl.deblock_frame(rec, *recon)
        yield f"deblock_{tag}", np.concatenate([a.ravel() for a in db])
        t, *sao = impl.sao_encode_frame(_planes(f, f"f{cur}"), db, impl.lam(q))
        yield f"sao_{tag}_table", np.frombuffer(t.tobytes(), np.uint8)
        yield f"sao_{tag}_planes", np.concatenate([a.ravel() for a in sao])

    dl = _load("deblock.npz")
    yield "deblock_lines", np.array([impl.deblock_line(v, int(q), int(st), bool(lu)) for v, q, st, lu
                                     in zip(dl["line_in"], dl["line_qp"], dl["line_strength"],
                                            dl["line_luma"])])
    sa = _load("sao.npz")
    yield "sao_filter", np.concatenate([a.ravel() for a in impl.sao_filter_frame(
        sa["filter_table"].view(SAO_DTYPE), *_planes(sa, "filter_in"))])


def run_child(build, which, out):
    sys.path.insert(0, ROOT)
    from oracle.oracle import Oracle, Reference
    path = BUILDS[build][which]
    impl = Reference(path) if which == "ref" else Oracle(path)
    np.savez(out, **dict(workloads(impl, which == "ref")))


_UB_LINE = re.compile(r"^(?P<file>[^:\s]+):(?P<line>\d+):(?P<col>\d+): runtime error: (?P<msg>.*)$")


def run(build, which, tmp):
    """Run the workloads through one build in a child process; returns the
    outputs and, for UBSan, {(file, line): [messages]}."""
    out = os.path.join(tmp, f"{build}_{which}.npz")
    env = dict(os.environ)
    log = os.path.join(tmp, f"ub_{which}")
    env["UBSAN_OPTIONS"] = f"log_path={log}:print_stacktrace=0:halt_on_error=0"
    subprocess.run([sys.executable, "-m", "oracle.ub_check", "run", build, which, out],
                   cwd=ROOT, env=env, check=True, timeout=600)
    sites = collections.defaultdict(set)
    if build == "ubsan":
        for name in os.listdir(tmp):
            if not name.startswith(f"ub_{which}."):
                continue
            with open(os.path.join(tmp, name)) as f:
                for line in f:
                    mt = _UB_LINE.match(line.strip())
                    if mt:
                        fn = mt["file"]
                        for pre in ("/root/reference/src/", HERE + "/"):
                            if fn.startswith(pre):
                                fn = fn[len(pre):]
                        msg = re.sub(r"-?\d+", "N", mt["msg"])
                        sites[(fn, int(mt["line"]))].add(msg)
    return dict(np.load(out)), {k: sorted(v) for k, v in sorted(sites.items())}


def differing(a, b):
    """Keys of outputs both runs produced whose values differ."""
    return sorted(k for k in a if k in b and not np.array_equal(a[k], b[k]))


def report():
    with tempfile.TemporaryDirectory() as tmp:
        res = {}
        for build in ("O3", "ubsan", "fwrapv"):
            for which in ("oracle", "ref"):
                res[build, which] = run(build, which, tmp)
    lines = []
    for which in ("oracle", "ref"):
        lines.append(f"UBSan sites, {which}:")
        for (fn, ln), msgs in res["ubsan", which][1].items():
            lines.append(f"  {fn}:{ln}: {'; '.join(msgs)}")
        lines.append(f"-fwrapv vs -O3 outputs differing, {which}: "
                     f"{differing(res['O3', which][0], res['fwrapv', which][0])}")
        lines.append(f"UBSan (-O1) vs -O3 outputs differing, {which}: "
                     f"{differing(res['O3', which][0], res['ubsan', which][0])}")
    lines.append(f"oracle vs reference (-O3) outputs differing: "
                 f"{[k for k in differing(res['O3', 'oracle'][0], res['O3', 'ref'][0])]}")
    return res, "\n".join(lines)


if __name__ == "__main__":
    if sys.argv[1] == "run":
        run_child(*sys.argv[2:5])
    else:
        print(report()[1])
