"""hsolve command line (apps/cli.cpp:116-148, 250-465): read matrices (+ config), solve on the
GPU, write the results.

    python -m hsolve_amd.cli --A A.npz [--B B.npz] [--config config.json] --out DIR [overrides...]

Outputs in DIR: eigenvalues.npy (m), eigenvectors.npy (n x m, Fortran order; complex128 when an
input has a complex dtype), complex_eigenvalues.npy / complex_eigenvectors.npy only when there are
genuinely complex eigenvalues (never for the definite pencils this engine solves; stale ones are
removed), report.json (report.json schema, with the input notes as "input: ..." warnings).
Overrides replace the configuration's fields (--sigma sets nearest mode, --interval / --smallest /
--largest the window modes) and the merged configuration is validated again. A summary goes to
stdout unless -q. Exit status: 0 = all pairs accepted and the certificate passed, 1 = finished
but not successful, 2 = error (message on stderr as "hsolve: error [code]: message").
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

from . import config as cfgmod
from .errors import HsolveError, InvalidInputError


def _parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="hsolve", description=__doc__.split("\n\n")[0],
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--A", dest="a", help="matrix A (.npz scipy sparse, .npy dense, .mtx); required except with --version "
                                          "or --case")
    ap.add_argument("--case", help="a case of the reference's data layout (a directory with A.npz, optional B.npz, "
                                   "config.json; or a case name under ROOT/cases): its matrices and configuration "
                                   "(command-line overrides apply on top)")
    ap.add_argument("--root", help="data root of --case names (default: $HS_DATA_ROOT or ./data)")
    ap.add_argument("--B", dest="b", help="matrix B (generalized problem)")
    ap.add_argument("--config", help="config.json")
    ap.add_argument("--out", default="hsolve-out", help="output directory (default: %(default)s)")
    ap.add_argument("--sigma", type=float, help="target shift (nearest mode)")
    ap.add_argument("--k", type=int, help="number of eigenpairs (nearest mode)")
    win = ap.add_mutually_exclusive_group()
    win.add_argument("--interval", type=float, nargs=2, metavar=("LO", "HI"), help="all eigenvalues in [LO, HI]")
    win.add_argument("--smallest", type=int, help="k algebraically smallest")
    win.add_argument("--largest", type=int, help="k algebraically largest")
    ap.add_argument("--tol", type=float, help="backward-error tolerance")
    ap.add_argument("--threads", type=int, help="host threads (0 = all available)")
    ap.add_argument("--precision", choices=("auto", "double", "single"), help="working factor precision")
    ap.add_argument("--slice-size", type=int, help="max eigenvalues per shift in sliced targets (0 = auto)")
    ap.add_argument("--block-size", type=int, help="Krylov block size (0 = auto)")
    ap.add_argument("--storage", choices=("auto", "full", "upper", "lower"), default="auto",
                    help="how the files store the matrices (default: detect one stored triangle)")
    ap.add_argument("-v", "--verbose", action="store_true", help="debug output")
    ap.add_argument("-q", "--quiet", action="store_true", help="no summary on stdout")
    ap.add_argument("--version", action="store_true", help="print the version")
    return ap


def merged_config(args) -> dict:
    cfg = cfgmod.parse_config(json.loads(open(args.config, encoding="utf-8").read()) if args.config else {})
    t = {}
    if args.sigma is not None:
        t["mode"], t["sigma"] = "nearest", args.sigma
    if args.k is not None:
        t["k"] = args.k
    if args.interval is not None:
        t["mode"], t["lower"], t["upper"] = "interval", args.interval[0], args.interval[1]
    if args.smallest is not None:
        t["mode"], t["k"] = "smallest", args.smallest
    if args.largest is not None:
        t["mode"], t["k"] = "largest", args.largest
    over = {"target": t}
    if args.tol is not None:
        over["tol"] = args.tol
    if args.threads is not None:
        over["threads"] = args.threads
    if args.precision is not None:
        over["factor_precision"] = args.precision
    if args.slice_size is not None:
        over["slice_size"] = args.slice_size
    if args.block_size is not None:
        over["block_size"] = args.block_size
    # re-validate the merged configuration (apps/cli.cpp:386)
    return cfgmod.parse_config(json.loads(cfgmod.config_to_json(cfgmod.parse_config(over, cfg))))


def write_outputs(out: str, res) -> None:
    from .io import write_npy_matrix, write_npy_vector
    os.makedirs(out, exist_ok=True)
    write_npy_vector(os.path.join(out, "eigenvalues.npy"), np.asarray(res.eigenvalues, dtype=np.float64))
    write_npy_matrix(os.path.join(out, "eigenvectors.npy"), res.eigenvectors)
    cev, cvec = os.path.join(out, "complex_eigenvalues.npy"), os.path.join(out, "complex_eigenvectors.npy")
    if res.complex_eigenvalues.size:
        write_npy_vector(cev, res.complex_eigenvalues)
        write_npy_matrix(cvec, res.complex_eigenvectors)
    else:
        for f in (cev, cvec):
            if os.path.exists(f):
                os.remove(f)
    with open(os.path.join(out, "report.json"), "w", encoding="utf-8") as f:
        f.write(cfgmod.report_to_json(res.report) + "\n")


def print_summary(r: dict) -> None:
    print(f"n={r.get('n')}  problem={r.get('problem')}  time={r.get('time_total_s', 0.0):.3f}s  "
          f"peak_rss={r.get('peak_rss_bytes', 0) / 2 ** 30:.2f} GiB  success={'yes' if r.get('success') else 'no'}")
    print(f"{'#':>5} {'Re(lambda)':>24} {'Im(lambda)':>24} {'backward_err':>12} {'ok':>4}")
    for i, p in enumerate(r.get("pairs", [])):
        print(f"{i:5d} {p['lambda_re']:24.16e} {p['lambda_im']:24.16e} {p['backward_error']:12.3e} "
              f"{'yes' if p['accepted'] else 'NO':>4}")
    c = r.get("certificate", {})
    if c.get("performed") and c.get("count_inertia", -1) >= 0:
        verdict = ("passed" if c.get("rigorous") else "passed (screening only)") if c.get("passed") else "FAILED"
        print(f"certificate ({c.get('method')}): [{c.get('lower'):g}, {c.get('upper'):g}] "
              f"inertia={c.get('count_inertia')} found={c.get('count_found')} -> {verdict}")
    for w in r.get("warnings", []):
        print(f"warning: {w}")


def main(argv=None) -> int:
    # concurrent spectrum slices need distinct hardware queues (HIP's default is 4); read when HIP
    # initializes, so it applies when this command owns the process (HSG_HW_QUEUES overrides 8)
    os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("HSG_HW_QUEUES", "8")
    ap = _parser()
    try:
        args = ap.parse_args(argv)
    except SystemExit as e:
        return 0 if e.code == 0 else 2
    if args.version:
        print(f"hsolve_amd {cfgmod.VERSION} (MI355X engine)")
        return 0
    if args.case:
        from .case import _case_dir
        d = _case_dir(args.case, args.root)
        if not (d / "A.npz").exists():
            print(f"hsolve: error [io]: no case at {d}", file=sys.stderr)
            return 2
        args.a = str(d / "A.npz")
        if args.b is None and (d / "B.npz").exists():
            args.b = str(d / "B.npz")
        if args.config is None and (d / "config.json").exists():
            args.config = str(d / "config.json")
    if not args.a:
        print("hsolve: error [invalid_input]: --A is required", file=sys.stderr)
        return 2
    try:
        cfg = merged_config(args)
        from .api import solve
        os.makedirs(args.out, exist_ok=True)
        res = solve(args.a, args.b, config=cfg, storage=args.storage, verbose=1 if args.verbose else 0)
        write_outputs(args.out, res)
        if not args.quiet:
            print_summary(res.report)
        return 0 if res.success else 1
    except HsolveError as e:
        print(f"hsolve: error [{e.code}]: {e}", file=sys.stderr)
        return 2
    except (OSError, ValueError) as e:   # unreadable config / output directory, bad JSON
        print(f"hsolve: error [{InvalidInputError.code if isinstance(e, ValueError) else 'io'}]: {e}", file=sys.stderr)
        return 2
    except Exception as e:  # e.g. no GPU: the engine has no CPU path
        print(f"hsolve: error [internal]: {e}", file=sys.stderr)
        return 2


if __name__ == "__main__":
    import warnings
    from .errors import HsolveWarning
    warnings.simplefilter("ignore", HsolveWarning)   # the exit status carries success
    sys.exit(main())
