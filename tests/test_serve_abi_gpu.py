"""The request server through its C ABI (mjk_server_create, mjk_submit,
mjk_poll, mjk_wait_any, mjk_result):

* J/K requests return what a direct build of the same density returns, and
  RHF requests the device-resident RHF's energy and density (the RHF runs
  mjk_rhf_* inside the server);
* admission on tracked bytes: with a budget that holds one request at a
  time, no two requests run together and the held bytes stay within the
  budget; with the default budget, requests overlap;
* a request whose estimate exceeds the budget is refused at submission, and
  malformed requests are refused with their reason.
"""
import ctypes
import time

import numpy
import pytest

from pyscf_amd import gto, lib, workloads
from pyscf_amd.scf import _vhf, hf, rhf
from pyscf_amd.server import JKServer, Request

pytestmark = pytest.mark.gpu


def chain(i, n=4, basis='def2-svp'):
    return gto.M(atom=workloads.random_chain([6, i], nmin=n, nmax=n), basis=basis)


def density(mol, seed):
    c = numpy.random.default_rng(seed).standard_normal((mol.nao_nr(), max(2, mol.nelectron // 2))) / 4
    return c @ c.T


def test_results_equal_direct_paths():
    mols = [chain(i) for i in range(4)] + [chain(9, n=3, basis='def2-tzvp')]
    srv = JKServer(nslots=4, conv_tol=1e-10)
    futs = []
    for i, m in enumerate(mols):
        kind = 'rhf' if i % 2 else 'get_jk'
        req = Request(kind, m, density(m, i) if kind == 'get_jk' else None)
        futs.append((req, srv.submit(req)))
    for (req, f), m in zip(futs, mols):
        out = f.result(timeout=300)
        print(req.kind, m.nao_nr(), 'slot', req.info.slot, 'held peak %.1f MB' % (req.info.hbm_peak / 2**20),
              'estimate %.1f MB' % (req.info.hbm_estimate / 2**20), flush=True)
        assert req.info.state == 2 and req.t_arrival <= req.t_start <= req.t_done
        if req.kind == 'get_jk':
            opt = _vhf.VHFOpt(m, direct_scf_tol=1e-13)
            j0, k0 = hf.get_jk(m, req.dm, vhfopt=opt)
            assert abs(out[0] - j0).max() <= 1e-12 and abs(out[1] - k0).max() <= 1e-12
        else:
            mf = rhf.DeviceRHF(m, direct_scf_tol=1e-13)
            mf.conv_tol = 1e-10
            e = mf.kernel()
            assert out[1] and abs(out[0] - e) <= 1e-10
            assert abs(req.dm_out - mf.dm).max() <= 1e-8
            assert req.info.cycles == mf.cycles
    srv.close()


def test_admission_serialises_when_budget_holds_one():
    mols = [chain(20)] * 4
    probe = JKServer(nslots=1)
    probe.submit(Request('rhf', mols[0])).result(timeout=300)
    probe.close()
    est = max(r.info.hbm_estimate for r in probe.finished)
    peak = max(r.info.hbm_peak for r in probe.finished)
    budget = int(1.5 * max(est, peak))
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    lib.load_library().mjk_mem_info(0, ctypes.byref(free), ctypes.byref(total))
    srv = JKServer(nslots=4, hbm_fraction=budget / total.value)
    held = []
    futs = [srv.submit(Request('rhf', m)) for m in mols]
    while not all(f.done() for f in futs):
        held.append(srv.hbm_held())
        time.sleep(0.002)
    for f in futs:
        assert f.result()[1]
    srv.close()
    fin = sorted(srv.finished, key=lambda r: r.t_start)
    print('budget %.1f MB, max held %.1f MB' % (budget / 2**20, max(held) / 2**20), flush=True)
    assert max(held) <= srv.budget
    assert all(fin[i + 1].t_start >= fin[i].t_done for i in range(len(fin) - 1))


def test_refusals():
    srv = JKServer(nslots=1, hbm_fraction=1e-9)
    with pytest.raises(MemoryError, match='budget'):
        srv.submit(Request('rhf', chain(1))).result(timeout=10)
    srv.close()
    srv = JKServer(nslots=1)
    odd = gto.M(atom='H 0 0 0; H 0 0 0.74; H 0 0 1.6', basis='def2-svp', spin=1)
    with pytest.raises(ValueError, match='even electron count'):
        srv.submit(Request('rhf', odd)).result(timeout=10)
    srv.close()
