"""Shared helpers: move an oracle-built level into the GPU import layout, error norms."""
import numpy as np


def level_arrays(osim, level=0):
    """Arrays of one oracle-built level in the lbm3_import_level layout."""
    a = {n: osim.array(n, level) for n in ("shifter", "rowI", "rowJ", "rowK", "rowJPlus", "rowKPlus", "rowJKPlus",
                                          "wallMap", "wallIndex", "wallData", "linkLength", "periodicIndex",
                                          "periodicSource")}
    n_open = osim.open_count(level)
    idx = [osim.array("open%dIndex" % b, level) for b in range(n_open)]
    src = [osim.array("open%dSource" % b, level) for b in range(n_open)]
    a["openCount"] = np.array([len(x) for x in idx], np.int32)
    a["openIndex"] = np.concatenate(idx).astype(np.int32) if idx else np.zeros(0, np.int32)
    a["openSource"] = np.concatenate(src).astype(np.int32) if src else np.zeros(0, np.int32)
    a["wallCount"] = np.array(int((a["wallMap"] == -3).sum()))
    if level < osim.levels - 1:
        for n in ("f2cIndex", "f2cChild", "c2fIndex", "c2fChild", "leftoverIndex", "leftoverParent", "leftoverNbrI",
                  "leftoverNbrJ", "leftoverNbrK"):
            a[n] = osim.array(n, level).astype(np.int32)
    return a


def rel_max(a, b, scale=None):
    """max |a - b| / max |b| (relative max-norm)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / s)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
