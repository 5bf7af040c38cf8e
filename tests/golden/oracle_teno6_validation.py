"""Runs the reference's TENO6 validation case (128^3, 500 iterations) through
the CPU oracle and prints the four post-processing reductions next to the
values the reference's benchmark README holds
(apps/performance_benchmark/README.MD:46-47).  Takes ~20 minutes on 8 cores;
its output is kept in oracle_teno6_validation.txt beside this script."""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from opensbli_amd import native as nat, problems as P  # noqa: E402

README = {"KE": 2.65363e+05, "dilatation_dissipation": 1.16991e+05,
          "enstrophy_dissipation": 1.79666e+06, "rhom": 2.09715e+06}


def main():
    lib = nat.Library(os.path.join(ROOT, "oracle", "libosb_oracle.so"))
    pr = P.tgv_teno6_validation()
    t0 = time.time()
    with lib.context(pr.desc) as c:
        c.set_state(pr.q0)
        for k in range(10):
            c.step(pr.steps // 10)
            print(f"iter {c.stats().iter} elapsed {time.time() - t0:.0f}s", flush=True)
        st = c.stats()
        diag = c.tgv_diagnostics()
    print(f"{st.iter}, {st.time:.5e}, " + ", ".join(f"{v:.5e}" for v in diag))
    for (k, ref), v in zip(README.items(), diag):
        print(f"{k}: oracle {v:.6e} reference {ref:.5e} rel {abs(v - ref) / ref:.2e}")


if __name__ == "__main__":
    main()
