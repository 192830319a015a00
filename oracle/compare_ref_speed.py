"""Speed of the oracle's CPU counting paths against the reference's own OpenMP
programs, both on this host with the same thread count.

    python oracle/compare_ref_speed.py [--workload 3] [--threads N]

Runs only where the reference tree is present (the build container): `make
-C oracle ref` compiles the reference programs from their sources into
oracle/_ref/ (git-ignored), the workload is written in the reference text
format to /tmp, and each program's phase lines are compared with the oracle's
timings of the same phases (the oracle restates
Multicore Implementations/tc_merge_sign_enu.cpp and tc_hash_sign_enu.cpp).
The oracle is what bench.py times as the CPU baseline on the GPU host; this
script shows that it runs at the reference's own speed.  Nothing here runs
on a GPU machine.
"""
import argparse
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def phases(text):
    out = {}
    for m in re.finditer(r"\[(.+?)\] took ([0-9.]+) seconds", text):
        out[m.group(1)] = float(m.group(2))
    b = re.search(r"Balanced\s*=\s*(\d+)", text)
    u = re.search(r"Unbalanced\s*=\s*(\d+)", text)
    out["counts"] = (int(b.group(1)), int(u.group(1))) if b and u else None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", HERE, "ref", "liboracle.so"], check=True)
    import signed_triangle_counting_amd as stc
    g = stc.gen_workload(args.workload)
    path = "/tmp/stc_compare_w%d.txt" % args.workload
    if not os.path.exists(path):
        g.write(path)
    env = dict(os.environ, OMP_NUM_THREADS=str(args.threads))
    rows = []
    for prog in ("tc_merge", "tc_hash"):
        exe = oracle.reference_binary(prog)
        r = subprocess.run([exe, path], env=env, capture_output=True, text=True, check=True)
        rows.append((prog, phases(r.stdout)))
    t0 = time.time()
    csr = oracle.Csr(g.n, g.u, g.v, g.sign, args.threads)
    build = time.time() - t0
    tm, b, u = min((csr.time_merge(args.threads) for _ in range(3)), key=lambda x: x[0])
    th, hb, b2, u2 = min((csr.time_hash(args.threads) for _ in range(3)), key=lambda x: x[0])
    print("| program | threads | CSR build (s) | preprocessing (s) | counting (s) | counts |")
    print("|---|---|---|---|---|---|")
    for prog, p in rows:
        pre = [v for k, v in p.items() if k.startswith("Algorithm-Specific Preprocessing")]
        print("| reference `%s` | %d | %.3f | %s | %.3f | %s |" % (
            prog, args.threads, p.get("Build Forward CSR", float("nan")),
            "%.3f" % pre[0] if pre else "—",
            p.get("Triangle Counting Execution", float("nan")), p["counts"]))
    print("| oracle merge (port) | %d | %.3f | — | %.3f | %s |" % (args.threads, build, tm, (b, u)))
    print("| oracle hash (port) | %d | %.3f | %.3f | %.3f | %s |" % (args.threads, build, hb, th,
                                                                    (b2, u2)))
    print("oriented edges: %d; best of 3 runs for the oracle, one run of each reference program"
          % csr.m)


if __name__ == "__main__":
    main()
