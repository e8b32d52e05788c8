"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import PKG, ROOT


def _declared():
    src = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mvr_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(os.path.join(PKG, "lib", "libmvr.so"))
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_header():
    import mvr._capi as capi
    assert set(_declared()) == set(capi.SIGNATURES)


def test_plaid_index_rejects_inconsistent_ivf():
    """mvr_plaid_index_create checks, before touching a device, that IVF[c]
    lists exactly the docs holding code c (the search depends on it)."""
    import numpy as np
    import mvr._capi as capi
    L = capi.lib()
    doclens = np.array([2, 1], np.int64)
    codes = np.array([0, 1, 1], np.int32)
    res = np.zeros((3, 32), np.uint8)
    cent = np.zeros((2, 128), np.float32)
    bw = np.zeros(4, np.float32)
    for ivf, lens in ((np.array([0, 0, 1], np.int32), np.array([1, 2], np.int64)),      # right
                      (np.array([0, 1], np.int32), np.array([1, 1], np.int64)),         # doc 0 missing in c=1
                      (np.array([0, 1, 0], np.int32), np.array([1, 2], np.int64))):     # not ascending
        d = capi.PlaidIndexDesc(n_doc=2, n_centroid=2, dim=128, nbits=2, n_tok=3, n_ivf=ivf.size,
                                doclens=doclens.ctypes.data, codes=codes.ctypes.data, residuals=res.ctypes.data,
                                centroids=cent.ctypes.data, bucket_weights=bw.ctypes.data, ivf=ivf.ctypes.data,
                                ivf_lengths=lens.ctypes.data)
        h = ctypes.c_void_p()
        st = L.mvr_plaid_index_create(ctypes.byref(d), 0, ctypes.byref(h))
        if ivf[1] == 0 and ivf.size == 3:
            assert st != 1, L.mvr_plaid_last_error()   # valid: fails only for want of a device here
            if st == 0:
                L.mvr_plaid_index_free(h)
        else:
            assert st == 1, (st, L.mvr_plaid_last_error())
            assert b"ivf list" in L.mvr_plaid_last_error()
