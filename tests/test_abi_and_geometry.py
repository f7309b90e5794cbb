"""CPU-side checks of the C ABI: the library loads, exports every entry point that
include/clm_hip.h declares, and its host Box construction equals the oracle's bit
for bit (the device binning consumes these numbers)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "clm_hip.h")).read()
    return sorted(set(re.findall(r"\b(clm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("clm_create", "clm_set_box", "clm_set_positions", "clm_update_cell_list", "clm_neighborlist",
              "clm_energy_forces", "clm_energy_adjoint", "clm_histogram", "clm_min_distance"):
        assert s in syms


def test_library_exports_every_declared_symbol(clm):
    L = clm.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(declared_symbols()) <= set(clm._lib.EXPORTS) | set(declared_symbols())
    assert L.clm_version().decode().startswith("celllistmap-hip")


def test_create_without_gpu_fails_cleanly(clm):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    p = C.c_void_p()
    rc = clm.lib().clm_create(C.byref(p), 0, 0, 3)
    assert rc != 0 and not p.value


def _lib_box(clm, dtype, uc, rc, lcell, celltype):
    info = clm._lib.BoxInfo()
    uc = np.asarray(uc, np.float64)
    dim = uc.shape[0]
    cell = np.ascontiguousarray((np.diag(uc) if uc.ndim == 1 else uc).T.reshape(-1))
    code = clm.lib().clm_box_compute(dtype, dim, cell.ctypes.data_as(C.c_void_p), celltype, float(rc), lcell,
                                      None, C.byref(info))
    return code, info


@pytest.mark.parametrize("dtype", [0, 1])
def test_box_matches_oracle_bitwise(clm, O, dtype):
    rng = np.random.default_rng(20 + dtype)
    npdt = np.float32 if dtype == 0 else np.float64
    checked = 0
    while checked < 300:
        dim = int(rng.integers(2, 4))
        tric = bool(rng.integers(0, 2))
        lcell = int(rng.integers(1, 4))
        rc = float(npdt(1 + 3 * rng.random()))
        if tric:
            uc = (-30 + 60 * rng.random((dim, dim))).astype(npdt).astype(np.float64)
        else:
            uc = (5 + 40 * rng.random(dim)).astype(npdt).astype(np.float64)
        code, info = _lib_box(clm, dtype, uc, rc, lcell, 1 if tric else 0)
        try:
            ob = O.make_box(uc, rc, lcell=lcell, dtype=npdt)
        except O.OracleError as e:
            assert code == 3 and e.code == 2   # both reject the cell
            continue
        assert code == 0
        for f in ("cutoff", "cutoff_sqr"):
            assert getattr(info, f) == getattr(ob, f)
        assert list(info.nc) == list(ob.nc)
        for a, b in (("cell_size", "cs"), ("cbmin", "cbmin"), ("cbmax", "cbmax")):
            assert list(getattr(info, a)) == list(getattr(ob, b)), a
        for a, b in (("aligned", "aligned"), ("rotation", "rot"), ("input", "input")):
            assert list(getattr(info, a)) == list(getattr(ob, b)), a
        checked += 1


def test_box_errors(clm):
    assert _lib_box(clm, 1, [10.0, 10.0, 10.0], 5.0, 1, 0)[0] == 3   # plane distance == 2 rc
    assert _lib_box(clm, 1, [10.0, 10.0, 10.0], 1.0, 0, 0)[0] == 1   # lcell < 1
    assert _lib_box(clm, 1, [10.0, 10.0, 10.0, 10.0], 1.0, 1, 0)[0] == 2  # dimension 4


def test_python_box_info_documented(clm):
    b = clm.box_info(np.array([[100, 50, 0], [0, 120, 0], [0, 0, 130]]), 10.0)
    assert b["nc"] == [20, 13, 16]
    assert clm.box_info([10, 10, 10], 1.0)["nc"] == [13, 13, 13]
    with pytest.raises(clm.ArgumentError):
        clm.box_info([10, 10, 10], 5.0)


def test_struct_layouts_match_header(clm, tmp_path):
    """The ctypes mirrors of clm_softcore and clm_box_info have the C layout: a few lines
    of C compiled against include/clm_hip.h print sizes and field offsets."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text("""#include <stddef.h>
#include <stdio.h>
#include "clm_hip.h"
int main(void) {
    printf("%zu %zu %zu %zu\\n", sizeof(clm_softcore), offsetof(clm_softcore, eps), offsetof(clm_softcore, sigma),
           offsetof(clm_softcore, form));
    printf("%zu %zu %zu %zu\\n", sizeof(clm_box_info), offsetof(clm_box_info, cutoff), offsetof(clm_box_info, input),
           offsetof(clm_box_info, origin));
    return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sc = [int(v) for v in out[0].split()]
    S = clm._lib.SoftCore
    assert sc == [C.sizeof(S), S.eps.offset, S.sigma.offset, S.form.offset]
    bi = [int(v) for v in out[1].split()]
    B = clm._lib.BoxInfo
    assert bi == [C.sizeof(B), B.cutoff.offset, B.input.offset, B.origin.offset]
