"""Comparison helpers shared by the parity tests."""
import numpy as np


def canon_keys(origins: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(origins, dtype=np.int64)


def voxel_points(origins, li, off, h):
    vox = origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    return vox.astype(np.float32) * np.float32(h)


def on_surface(d, pts, h, sign_eps_vox=1e-6):
    """Voxels whose |d| is below 1e-6 voxel or below what f32 coordinates
    resolve (16 ulps of the largest coordinate): their sign is not defined."""
    ulp = np.spacing(np.max(np.abs(pts), axis=1).astype(np.float32)).astype(np.float64)
    return np.abs(d) <= np.maximum(sign_eps_vox * h, 16.0 * ulp)


def compare_volumes(gpu, ref, voxel_size: float, tol_vox: float = 1e-4, sign_eps_vox: float = 1e-6, signed=True):
    """Bit-exact leaf keys and masks; |dv| <= tol voxels on active voxels; signs
    bit-exact wherever the voxel is not on the surface (see on_surface).
    Returns an error summary."""
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
        li, off = np.nonzero(act)
        big = ~on_surface(r, voxel_points(ref.origins, li, off, voxel_size), voxel_size, sign_eps_vox)
        mism = np.signbit(g[big]) != np.signbit(r[big])
        out["sign_mismatches"] = int(mism.sum())
        out["on_surface_exempt"] = int((~big).sum())
    return out
