"""SaveResults (SourceCode/SaveResults.c:32-92) on device-resident results,
both layouts, every written file read back and compared bit for bit with the
matrices it holds: S_hat, then U_hat_k<j>_allFreqs / V_hat_k<j>_allFreqs (opt
1, 1-based) or U_hat_Freq<m>_allK / V_hat_Freq<m>_allK (opt 2, 0-based).  The
large case has more rows than one 64 MB slice of the streamed writer."""
import ctypes
import os

import numpy as np
import pytest

import petsc_bin
from conftest import crandn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("opt", [1, 2])
@pytest.mark.parametrize("N,k,nwe", [(3000, 5, 7), (2_200_000, 2, 2)])
def test_save_results_files(pkg, tmp_path, opt, N, k, nwe):
    print(f"SaveResults opt={opt} N={N} k={k} Nw_eff={nwe}", flush=True)
    L = pkg.lib()
    assert L.RSVDT_Initialize(None, None) == 0
    rng = np.random.default_rng(N + opt)
    U = crandn(rng, N, k * nwe)
    Vm = crandn(rng, N, k * nwe)
    S = rng.random((k, nwe))
    res = pkg.Resolvent_matrices()
    res.U_hat = pkg.mat_from_dense(U, k).value
    res.V_hat = pkg.mat_from_dense(Vm, k).value
    res.S_hat = pkg.mat_from_dense(S.astype(np.complex128)).value
    V = pkg.RSVDt_vars()
    V.RSVD.k, V.RSVD.Nw_eff, V.SaveResultsOpt = k, nwe, opt
    d = pkg.Directories()
    folder = str(tmp_path) + "/"
    d.FolderDir = folder.encode()
    try:
        assert L.SaveResults(ctypes.byref(res), ctypes.byref(V), ctypes.byref(d)) == 0
    finally:
        for h in (res.U_hat, res.V_hat, res.S_hat):
            pkg.destroy(h)
    assert np.array_equal(petsc_bin.read_dense(folder + "S_hat"), S.astype(np.complex128))
    files = sorted(os.listdir(folder))
    if opt == 1:
        assert len(files) == 1 + 2 * k
        for j in range(k):
            for name, X in (("U", U), ("V", Vm)):
                got = petsc_bin.read_dense(f"{folder}{name}_hat_k{j + 1}_allFreqs")
                assert got.shape == (N, nwe)
                assert np.array_equal(got, X[:, j::k])       # column m*k + j of every frequency m
    else:
        assert len(files) == 1 + 2 * nwe
        for m in range(nwe):
            for name, X in (("U", U), ("V", Vm)):
                got = petsc_bin.read_dense(f"{folder}{name}_hat_Freq{m}_allK")
                assert got.shape == (N, k)
                assert np.array_equal(got, X[:, m * k:(m + 1) * k])
