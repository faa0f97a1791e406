"""Comparison helpers shared by the parity tests."""
import numpy as np


def canon_keys(origins: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(origins, dtype=np.int64)


def compare_volumes(gpu, ref, voxel_size: float, tol_vox: float = 1e-4, sign_eps_vox: float = 1e-6, signed=True):
    """Bit-exact leaf keys and masks; |dv| <= tol voxels on active voxels; signs
    bit-exact wherever |d| > sign_eps voxels.  Returns an error summary."""
    assert gpu.origins.shape == ref.origins.shape, (gpu.origins.shape, ref.origins.shape)
    assert np.array_equal(gpu.origins, ref.origins), "leaf keys differ"
    assert np.array_equal(gpu.masks, ref.masks), "active masks differ"
    act = ref.active()
    g = gpu.values[act].astype(np.float64)
    r = ref.values[act].astype(np.float64)
    err = np.abs(g - r) / voxel_size
    out = {"n_active": int(act.sum()), "max_err_vox": float(err.max(initial=0.0)),
           "rms_err_vox": float(np.sqrt(np.mean(err ** 2))) if err.size else 0.0,
           "bit_exact_frac": float(np.mean(gpu.values[act] == ref.values[act])) if err.size else 1.0}
    if signed:
        big = np.abs(r) > sign_eps_vox * voxel_size
        mism = np.signbit(g[big]) != np.signbit(r[big])
        out["sign_mismatches"] = int(mism.sum())
    return out
