"""The C-ABI library loads and exports every entry point include/*.h declares; host-only
calls work without a device, and device calls fail loudly (no CPU fallback)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from powder6_amd import pt
from powder6_amd import problems as PR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        text = open(h).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        for m in re.finditer(r'^[A-Za-z_][\w \*]*?\b(pt_\w+)\s*\(', text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ('pt_plan_create', 'pt_plan_destroy', 'pt_costs', 'pt_batch', 'pt_eval', 'pt_last_error'):
        assert n in names


@pytest.mark.parametrize('name', declared_functions())
def test_library_exports(name):
    lib = ctypes.CDLL(pt.LIB_PATH)
    assert hasattr(lib, name), name


def test_abi_version():
    assert pt.lib().pt_abi_version() == 3


def test_phase_base_layout():
    pb = PR.pbso4_pnma()
    pb.y = np.ones_like(pb.tth)
    cp = pt.CProblem(pb)
    lib = pt.lib()
    # zero, disp, transp, 24 background slots, 18 instrument slots; a phase block is 8
    # cell / scale / po slots, 32 profile slots and 5 per atom
    assert lib.pt_phase_base(cp.ref(), 0) == 45
    assert lib.pt_phase_base(cp.ref(), 1) == len(pb.base) == 45 + 40 + 5 * len(pb.phases[0].atoms)


def test_compile_without_device_reports_counts():
    pb = PR.nacl_fm3m()
    pb.y = np.ones_like(pb.tth)
    info = pt.compile({'problem': pb})
    assert info['point_count'] == 5501 and info['nv'] == 12 and info['chains'] == 16
    assert info['stride'] == 3 * 12 + 12 and info['uniform_grid'] == 1
    assert info['reflection_count'] > 0


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason='checks the behaviour without a device')
def test_device_calls_fail_loudly_without_gpu():
    pb = PR.nacl_fm3m()
    pb.y = np.ones_like(pb.tth)
    with pytest.raises(pt.PTError, match='unavailable'):
        pt.Plan(pb)
    with pt.Plan(pb, device=False) as pl:
        with pytest.raises(pt.PTError, match='without a device'):
            pl.costs(pb.start_vector())


def test_multi_dataset_plan_needs_one_variable_list():
    """pt_plan_create_multi refuses datasets whose variable counts differ (checked before
    any device work)."""
    a = PR.nacl_fm3m()
    b = PR.nacl_fm3m()
    b.refine('p1.po')
    ca, cb = pt.CProblem(a), pt.CProblem(b)
    arr = (pt.PtProblem * 2)()
    arr[0], arr[1] = ca.struct, cb.struct
    h = ctypes.c_void_p()
    rc = pt.lib().pt_plan_create_multi(ctypes.cast(arr, ctypes.c_void_p), 2, ctypes.byref(h))
    assert rc < 0 and b'variable list' in pt.lib().pt_last_error()


def test_free_intensity_budget():
    """More than 2048 potential free reflection intensities in a dataset is refused with
    the reference's budget phrase (js/pt_gpu.js:175), before any device work (the
    molecular P2_1/c envelope holds 2,730 reflections)."""
    pb = PR.molecular_p21c()
    pb.phases[0].method = 'lebail'
    with pytest.raises(pt.PTError, match='free reflection intensities'):
        pt.Plan(pb, device=False)
