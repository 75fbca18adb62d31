"""Batched SCF and the continuous-batching server: every molecule's result
is the same alone and in the batch / in the mix (within tolerance)."""
import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf, batch, hf, rhf

pytestmark = pytest.mark.gpu


def chains(n, basis='def2-svp', nmax=5):
    return [gto.M(atom=workloads.random_chain([4, i], nmin=3, nmax=nmax), basis=basis) for i in range(n)]


def test_batched_jk_equals_single():
    mols = chains(5)
    b = batch.BatchRHF(mols)
    b.basis = b.engine.upload_batch(mols, 1e-13)
    dms = []
    for m in mols:
        c = numpy.random.default_rng(m.nao_nr()).standard_normal((m.nao_nr(), 6)) / 3
        dms.append(c @ c.T)
    js, ks = b.get_jk_batch(dms)
    for m, d, j, k in zip(mols, dms, js, ks):
        opt = _vhf.VHFOpt(m, direct_scf_tol=1e-13)
        j0, k0 = hf.get_jk(m, d, vhfopt=opt)
        assert abs(j - j0).max() < 1e-11 and abs(k - k0).max() < 1e-11


def test_batched_rhf_matches_single():
    mols = chains(4)
    b = batch.BatchRHF(mols)
    b.conv_tol = 1e-10
    e = b.kernel()
    assert b.converged.all()
    for m, ei in zip(mols, e):
        mf = rhf.RHF(m)
        mf.conv_tol = 1e-10
        assert abs(mf.kernel() - ei) < 1e-8


def test_server_alone_vs_mix():
    from pyscf_amd.server import JKServer, Request
    mols = chains(6)
    reqs = []
    for i, m in enumerate(mols):
        if i % 2:
            reqs.append(('rhf', m, None))
        else:
            c = numpy.random.default_rng(i).standard_normal((m.nao_nr(), 5)) / 3
            reqs.append(('get_jk', m, c @ c.T))
    srv = JKServer(nslots=4)
    alone = []
    for kind, m, dm in reqs:
        alone.append(srv.submit(Request(kind, m, dm)).result(timeout=120))
    futs = [srv.submit(Request(kind, m, dm)) for kind, m, dm in reqs]
    mix = [f.result(timeout=120) for f in futs]
    srv.close()
    for (kind, _, _), a, b in zip(reqs, alone, mix):
        if kind == 'get_jk':
            assert abs(a[0] - b[0]).max() < 1e-12 and abs(a[1] - b[1]).max() < 1e-12
        else:
            assert a[1] and b[1] and abs(a[0] - b[0]) < 1e-10
    # requests overlapped in the mix
    fin = sorted(srv.finished[len(reqs):], key=lambda r: r.t_start)
    assert any(fin[i + 1].t_start < fin[i].t_done for i in range(len(fin) - 1))


def test_server_refuses_oversized_request():
    from pyscf_amd.server import JKServer, Request
    srv = JKServer(nslots=1, hbm_fraction=1e-9)
    fut = srv.submit(Request('rhf', chains(1)[0]))
    with pytest.raises(MemoryError):
        fut.result(timeout=10)
    srv.close()
