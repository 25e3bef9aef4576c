"""Writes antitrap_2048_20000_vd<r>.npz: the CPU oracle's state after 20 000
steps of the 2048^2 coupled-T Ti-Nb configuration with the antitrapping
current and the solute-trapping factor switched on (a_t = 1/(2 sqrt 2),
V_D = Vp / r, i.e. V_p / V_D = r), seeded planar front (init_config(21)),
FP64. The state is reduced the same way as coupled_2048_20000.npz (full
rows, a window around the tip, every 8th coarse-T row, integral figures and
the tip row; see make_coupled_2048_20000.py). No pull-back falls inside the
run (step_pull_back = 29 936).

    python tests/golden/make_antitrap_2048_20000.py [ratio] [threads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "2d-rapid-solidification_amd"))

import oracle  # noqa: E402
from make_coupled_2048_20000 import N, SEED, STEPS, CHUNK, reduce_state  # noqa: E402
from rsolid import config_params  # noqa: E402

AT_COEF = 1.0 / (2.0 * np.sqrt(2.0))


def fixture_name(ratio: float) -> str:
    return f"antitrap_2048_20000_vd{ratio:g}.npz"


def main() -> None:
    ratio = float(sys.argv[1]) if len(sys.argv) > 1 else 2.0
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    oracle.set_threads(threads)
    p = config_params(N, True)
    p.at_coef = AT_COEF
    p.V_D = p.Vp / ratio
    st = oracle.State(p)
    st.init_config(SEED)
    assert st.constants().step_pull_back > STEPS
    m0 = st.solute_mass()
    t0 = time.time()
    done = 0
    while done < STEPS:
        st.step(CHUNK)
        done += CHUNK
        print("step", done, "elapsed %.0f s" % (time.time() - t0), flush=True)
    out = reduce_state(st)
    out.update({"N": np.int64(N), "seed": np.int64(SEED), "steps": np.int64(STEPS),
                "at_coef": np.float64(AT_COEF), "V_D": np.float64(p.V_D),
                "ratio": np.float64(ratio), "solute_mass0": np.float64(m0)})
    path = os.path.join(HERE, fixture_name(ratio))
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")
    st.close()


if __name__ == "__main__":
    main()
