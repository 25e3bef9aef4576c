"""Writes antitrap_2048_20000_vd<r>.npz: the CPU oracle's state after 20 000
steps of the 2048^2 coupled-T Ti-Nb configuration with the antitrapping
current and the solute-trapping factor switched on (a_t = 1/(2 sqrt 2),
V_D = Vp / r, i.e. V_p / V_D = r), seeded planar front (init_config(21)),
FP64. The state is reduced the same way as coupled_2048_20000.npz (full
rows, a window around the tip, every 8th coarse-T row, integral figures and
the tip row; see make_coupled_2048_20000.py). No pull-back falls inside the
run (step_pull_back = 29 936).

    python tests/golden/make_antitrap_2048_20000.py [ratio] [threads] [checkpoint.npz]

The run takes ~2 h on 8 cores. With a checkpoint path the oracle's full
state (psi, phi, c, coarse T, run constants) is saved there every 2 000
steps and a later call with the same path resumes from it; the resumed run
is bitwise the uninterrupted one (T_fine is rebuilt from the coarse T with
the i_left_end the last step used).
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
from rsolid.abi import RS_FIELD_C, RS_FIELD_PHI, RS_FIELD_PSI, RS_FIELD_T, RsConstants  # noqa: E402

AT_COEF = 1.0 / (2.0 * np.sqrt(2.0))
SAVE_EVERY = 2000


def fixture_name(ratio: float) -> str:
    return f"antitrap_2048_20000_vd{ratio:g}.npz"


def params(ratio: float):
    p = config_params(N, True)
    p.at_coef = AT_COEF
    p.V_D = p.Vp / ratio
    return p


def save_checkpoint(st, path: str, m0: float) -> None:
    k = st.constants()
    tmp = path + ".tmp.npz"
    np.savez(tmp, psi=st.get(RS_FIELD_PSI), phi=st.get(RS_FIELD_PHI), c=st.get(RS_FIELD_C),
             T=st.get(RS_FIELD_T), k=np.frombuffer(bytes(k), dtype=np.uint8), m0=np.float64(m0))
    os.replace(tmp, path)


def load_checkpoint(st, path: str) -> float:
    with np.load(path, allow_pickle=False) as z:
        st.put(RS_FIELD_PSI, z["psi"])
        st.put(RS_FIELD_PHI, z["phi"])
        st.put(RS_FIELD_C, z["c"])
        st.put(RS_FIELD_T, z["T"])
        k = RsConstants.from_buffer_copy(z["k"].tobytes())
        m0 = float(z["m0"])
    st.set_constants(k)
    st.interpolate_at(k.i_left_end_interp)
    return m0


def run(ratio: float, steps: int, ckpt: str | None, n: int = N) -> "oracle.State":
    p = params(ratio) if n == N else _small_params(ratio, n)
    st = oracle.State(p)
    st.init_config(SEED)
    m0 = st.solute_mass()
    if ckpt and os.path.exists(ckpt):
        m0 = load_checkpoint(st, ckpt)
        print("resumed at step", st.constants().step, flush=True)
    t0 = time.time()
    while st.constants().step < steps:
        todo = min(CHUNK, steps - st.constants().step)
        st.step(todo)
        done = st.constants().step
        print("step", done, "elapsed %.0f s" % (time.time() - t0), flush=True)
        if ckpt and done % SAVE_EVERY == 0 and done < steps:
            save_checkpoint(st, ckpt, m0)
    st.m0 = m0
    return st


def _small_params(ratio: float, n: int):
    p = config_params(n, True)
    p.at_coef = AT_COEF
    p.V_D = p.Vp / ratio
    return p


def main() -> None:
    ratio = float(sys.argv[1]) if len(sys.argv) > 1 else 2.0
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    ckpt = sys.argv[3] if len(sys.argv) > 3 else None
    oracle.set_threads(threads)
    st = run(ratio, STEPS, ckpt)
    assert st.constants().step_pull_back > STEPS
    out = reduce_state(st)
    out.update({"N": np.int64(N), "seed": np.int64(SEED), "steps": np.int64(STEPS),
                "at_coef": np.float64(AT_COEF), "V_D": np.float64(st.p.V_D),
                "ratio": np.float64(ratio), "solute_mass0": np.float64(st.m0)})
    path = os.path.join(HERE, fixture_name(ratio))
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")
    st.close()


if __name__ == "__main__":
    main()
