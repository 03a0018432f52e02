"""Writes the golden fixtures under tests/golden/ (run in a container that has the
reference sources; the fixtures are data and travel without them).

  coeff_ref_L{2,4,8}.npz  -- lambda_k and the Galerkin tables exactly as the reference's own
                             Coeff<double> (gpu_flow/code/util/coeff.cu, compiled unmodified
                             by oracle/Makefile into oracle/_ref/) computes them.
  data_state_kept.npz     -- the reference's bundled FP32 state (gpu_flow/data/*_00030000.bin,
                             257x512x2 complex64 theta/f/g and 1x1x2 F/G), reduced to the
                             modes the dealias mask keeps plus the full-array invariants.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/gpu_flow"
TWO = ["Iac", "Iab", "Ia1", "Ia2"]
THREE = ["Iac1", "Ibb1", "Iaab", "Iaad", "Iabc", "Ia11", "Ia12", "I111", "Ibbb", "Ib11", "Iaa1", "Iab2"]


def read_ref_dump(path):
    b = open(path, "rb").read()
    L = int(np.frombuffer(b[:4], np.int32)[0])
    off = 4
    lam = np.frombuffer(b[off:off + 8 * L], np.float64).copy()
    off += 8 * L
    out = {"lambda": lam}
    for n in TWO:
        out[n] = np.frombuffer(b[off:off + 8 * L * L], np.float64).reshape(L, L).T.copy()  # [i, j]
        off += 8 * L * L
    for n in THREE:
        out[n] = np.frombuffer(b[off:off + 8 * L ** 3], np.float64).reshape(L, L, L).transpose(2, 1, 0).copy()
        off += 8 * L ** 3
    return out


def read_state_file(path):
    """Reference binary format (output/matrix_folder_writer.cu:3-58)."""
    b = open(path, "rb").read()
    nmat = int(np.frombuffer(b[:4], np.int32)[0])
    off = 4
    mats = []
    for _ in range(nmat):
        nd = int(np.frombuffer(b[off:off + 4], np.int32)[0])
        off += 4
        dims = np.frombuffer(b[off:off + 4 * nd], np.int32).tolist()
        off += 4 * nd
        n = int(np.prod(dims))
        esz = (len(b) - off) // n if nmat == 1 else 8
        dt = np.complex64 if esz == 8 else np.complex128
        a = np.frombuffer(b[off:off + n * esz], dt).copy()
        off += n * esz
        mats.append((dims, a))
    return mats


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    dump = os.path.join(ROOT, "oracle", "_ref", "ref_coeff_dump")
    for L in (2, 4, 8):
        tmp = os.path.join(ROOT, "oracle", "_ref", f"coeff_L{L}.bin")
        subprocess.check_call([dump, str(L), tmp])
        np.savez_compressed(os.path.join(HERE, f"coeff_ref_L{L}.npz"), **read_ref_dump(tmp))
    # bundled state
    sys.path.insert(0, os.path.dirname(HERE))
    from helpers import kept_mask
    fields = {}
    inv = {}
    for name in ("theta", "f", "g", "F", "G"):
        (dims, a), = read_state_file(os.path.join(REF, "data", f"{name}_00030000.bin"))
        mox, N, L = dims
        arr = a.reshape(L, N, mox)
        fields[name + "_dims"] = np.array(dims, np.int32)
        if name in ("theta", "f", "g"):
            M = 2 * (mox - 1)
            keep = kept_mask(M, N)
            inv[name + "_removed_band_maxabs"] = np.abs(arr[:, ~keep]).max()
            fields[name] = arr[:, keep]  # (L, nkept) complex64, row-major order of kept (row, col)
        else:
            fields[name] = arr.reshape(-1)
    np.savez_compressed(os.path.join(HERE, "data_state_kept.npz"), **fields,
                        **{k: np.float64(v) for k, v in inv.items()})
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
