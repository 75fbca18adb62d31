"""Batched closed-shell RHF: many independent molecules converged together.

Every SCF iteration is ONE batched J/K build: one launch per quartet class
covers the shell pairs of all molecules (the molecules' pairs sit back to
back in the pair lists of ``mjk_basis_upload_batch``), and after the first
build each iteration replays the build's HIP graph.  Molecules that have
converged are marked inactive (``mjk_basis_set_active``): their blocks exit
at once, so an iteration costs what its active molecules cost.  Per-molecule
work (Fock assembly, CDIIS, generalized eigh) follows pyscf/scf/hf.py:47-242
as in ``rhf.RHF``.
"""

import time

import numpy

from pyscf_amd.scf import _vhf, rhf


class BatchRHF:
    def __init__(self, mols, direct_scf_tol=1e-13, fp32_tol=0.0, engine=None, log=None):
        self.mols = list(mols)
        self.log = log              # optional callable(str): one line per SCF cycle
        self.direct_scf_tol = direct_scf_tol
        self.fp32_tol = fp32_tol
        self.conv_tol = 1e-9
        self.max_cycle = 100
        self.engine = engine or _vhf.get_engine()
        self.basis = None
        self.e_tot = None
        self.converged = None
        self.cycles = 0
        self.build_times = []
        self.build_active = []      # active molecules of every build
        self.build_quartets = []    # evaluated quartets of every build (stats=True only)
        self.stats = False

    def _offsets(self):
        naos = [m.nao_nr() for m in self.mols]
        off = numpy.zeros(len(naos) + 1, dtype=numpy.int64)
        off[1:] = numpy.cumsum([n * n for n in naos])
        return naos, off

    def get_jk_batch(self, dms, active=None):
        """J and K for one density per molecule (list of (nao, nao)); with
        ``active`` (one flag per molecule) the inactive molecules are skipped
        and get zero J and K."""
        naos, off = self._offsets()
        if active is None:
            active = numpy.ones(len(naos), dtype=bool)
        active = numpy.asarray(active, dtype=bool)
        if not numpy.array_equal(active, getattr(self, '_active_set', None)):
            self.basis.set_active(active)
            self._active_set = active.copy()
        flat = numpy.zeros((1, int(off[-1])))
        for i in numpy.nonzero(active)[0]:
            flat[0, off[i]:off[i + 1]] = numpy.asarray(dms[i]).ravel()
        t0 = time.perf_counter()
        vj, vk = self.basis.get_jk(flat, 1, True, True, self.direct_scf_tol, self.fp32_tol, stats=self.stats)
        self.build_times.append(time.perf_counter() - t0)
        self.build_active.append(int(active.sum()))
        if self.stats:
            self.build_quartets.append(self.basis.last_stats['quartets'])
        js = [vj[0, off[i]:off[i + 1]].reshape(n, n) for i, n in enumerate(naos)]
        ks = [vk[0, off[i]:off[i + 1]].reshape(n, n) for i, n in enumerate(naos)]
        return js, ks

    def kernel(self):
        mols = self.mols
        if self.basis is None:
            self.basis = self.engine.upload_batch(mols, self.direct_scf_tol)
        drivers = [rhf.RHF(m) for m in mols]
        ints = [rhf.int1e_device(m, self.engine) for m in mols]   # on the device, one call per molecule
        h1e = [x['kin'] + x['nuc'] for x in ints]
        s1e = [x['ovlp'] for x in ints]
        enuc = [m.energy_nuc() for m in mols]
        dms = []
        for d, h, s in zip(drivers, h1e, s1e):
            e, c = d.eig(h, s)
            dms.append(d.make_rdm1(c, d.get_occ(e)))
        js, ks = self.get_jk_batch(dms)
        vhf = [j - .5 * k for j, k in zip(js, ks)]
        e_tot = [rhf.RHF.energy_elec(dm, h, v) + en for dm, h, v, en in zip(dms, h1e, vhf, enuc)]
        diis = [rhf.CDIIS(8) for _ in mols]
        active = numpy.ones(len(mols), dtype=bool)
        conv_g = numpy.sqrt(self.conv_tol)
        for cycle in range(self.max_cycle):
            new_dms = list(dms)
            for i in numpy.nonzero(active)[0]:
                f = diis[i].update(s1e[i], dms[i], h1e[i] + vhf[i])
                e, c = drivers[i].eig(f, s1e[i])
                occ = drivers[i].get_occ(e)
                new_dms[i] = drivers[i].make_rdm1(c, occ)
                drivers[i].mo_coeff, drivers[i].mo_occ = c, occ
            ddm = [new_dms[i] - dms[i] if active[i] else None for i in range(len(mols))]
            js, ks = self.get_jk_batch(ddm, active)
            self.cycles = cycle + 1
            for i in numpy.nonzero(active)[0]:
                vhf[i] = vhf[i] + js[i] - .5 * ks[i]
                e_last = e_tot[i]
                e_tot[i] = rhf.RHF.energy_elec(new_dms[i], h1e[i], vhf[i]) + enuc[i]
                fock = h1e[i] + vhf[i]
                g = drivers[i].get_grad(drivers[i].mo_coeff, drivers[i].mo_occ, fock)
                if abs(e_tot[i] - e_last) < self.conv_tol and numpy.linalg.norm(g) < conv_g:
                    active[i] = False
            dms = new_dms
            if self.log:
                self.log('cycle %d: %d of %d molecules active, J/K build %.1f ms'
                         % (cycle + 1, int(active.sum()), len(mols), 1e3 * self.build_times[-1]))
            if not active.any():
                break
        self.e_tot = numpy.array(e_tot)
        self.converged = ~active
        self.dms = dms
        return self.e_tot
