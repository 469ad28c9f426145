"""Shared helpers for the parity tests."""
import numpy as np

from robast_amd import _abi as A

FIELDS = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node")


def copy_soa(soa):
    return {k: np.array(v, copy=True) for k, v in soa.items()}


def compare(gpu, ora, pos_tol=1e-7, dir_tol=1e-9, t_tol=1e-15, check_nodes=True):
    """Per-ray parity: identical status / npoints / last node, terminal point
    within pos_tol cm (1e-7 cm = 1e-9 m), direction within dir_tol."""
    st = gpu["status"] == ora["status"]
    bad = np.nonzero(~st)[0]
    assert len(bad) == 0, f"{len(bad)} status mismatches, first {bad[:5]}: gpu {gpu['status'][bad[:5]]} oracle {ora['status'][bad[:5]]}"
    assert np.array_equal(gpu["npoints"], ora["npoints"])
    if check_nodes:
        assert np.array_equal(gpu["last_node"], ora["last_node"])
    dp = np.sqrt((gpu["x"] - ora["x"]) ** 2 + (gpu["y"] - ora["y"]) ** 2 + (gpu["z"] - ora["z"]) ** 2)
    dd = np.sqrt((gpu["dx"] - ora["dx"]) ** 2 + (gpu["dy"] - ora["dy"]) ** 2 + (gpu["dz"] - ora["dz"]) ** 2)
    assert dp.max() <= pos_tol, f"max |dpos| {dp.max():.3e} cm"
    assert dd.max() <= dir_tol, f"max |ddir| {dd.max():.3e}"
    fin = np.isfinite(ora["t"]) & (np.abs(ora["t"]) < 1e10)
    assert np.abs(gpu["t"] - ora["t"])[fin].max(initial=0) <= t_tol
    return dp.max()


def counts(status):
    return np.bincount(np.asarray(status), minlength=6)
