"""Many particles of uneven face counts solved together on the GPU.

``BEMRetBatch([cp1, cp2, ...])`` solves the retarded BEM problem of every
ComParticle over the same wavelengths in shared batched launches.  The
particles are grouped by face count; a group's matrices for a block of L
wavelengths form one batch of S x L matrices (S particles, particle-major,
the L wavelengths of one particle consecutive) that runs through the kernels
of BEMRet -- Green assembly, the four LU factorisations, the
many-right-hand-side solves, Delta^-1, the direct solve with FP64 refinement
and the extinction -- with the group's per-face geometry and refinement data
stacked per particle, so that each batch element reads its own particle
(geometry slots, ``gdiv`` = L in the C ABI).  The dense O(N^3) work of all
particles of one size shares every launch, which fills the GPU even for the
small particles; the groups follow each other on one stream.
``BatchSweepGraph`` captures one wavelength block of all groups as one HIP
graph and replays it for every block.

The reference solves particles one after another, one BEMRet per particle
and per wavelength (mnpbem/bem/bem_ret.py:293-525, 1569-1831); results per
particle equal that path's.
"""

import numpy as np
import torch

from . import _native as nat
from .bem import BEMRet, _par_tensor, needs_interchanges
from .greenfun import CompGreenRet


class _GroupGreen:
    """Stacked device geometry and refinement data of the particles of one
    face count: the CompGreenRet attributes that BEMRet uses, with geometry
    slots (``gdiv`` wavelengths per particle, pair offsets ``poff``)."""

    def __init__(self, greens):
        g0 = greens[0]
        self.greens = greens
        self.S = len(greens)
        self.n = g0.n
        self.device = g0.device
        self.order = g0.order
        if any(g.n != self.n or g.order != self.order for g in greens):
            raise ValueError("BEMRetBatch: a group needs one face count and one refinement order")
        self.nmat = max(g.nmat for g in greens)
        cat = lambda f: torch.cat([f(g) for g in greens]).contiguous()
        self.pos = cat(lambda g: g.pos)
        self.nvec = cat(lambda g: g.nvec)
        self.area = cat(lambda g: g.area)
        self.mat = [cat(lambda g, s=s: g.mat[s]) for s in (0, 1)]
        counts = np.array([g.npairs for g in greens], dtype=np.int64)
        self.poff = torch.as_tensor(np.concatenate([[0], np.cumsum(counts)])).to(self.device)
        self.npairs = int(counts.max())
        self.prow = cat(lambda g: g.prow)
        self.pcol = cat(lambda g: g.pcol)
        self.g = cat(lambda g: g.g)
        self.f = cat(lambda g: g.f)
        self.con = g0.con
        self.gdiv = 1

    def slot_lams(self, lams):
        """Batch layout of a wavelength block: element b = s L + l."""
        return np.tile(np.asarray(lams, dtype=float), self.S)

    def wavenumbers(self, enei_list):
        """(S L, nmat, 2) wavenumbers, element b from particle b // L."""
        L = self.gdiv
        out = np.zeros((len(enei_list), self.nmat, 2))
        for s, g in enumerate(self.greens):
            lam = enei_list[s * L:(s + 1) * L]
            k = np.array([[complex(e(x)[1]) for e in g.p1.eps] for x in lam])
            out[s * L:(s + 1) * L, :k.shape[1], 0] = k.real
            out[s * L:(s + 1) * L, :k.shape[1], 1] = k.imag
        return torch.as_tensor(out, dtype=torch.float64).to(self.device)

    def assemble(self, side, kmat, G, H, hsign, dtype, G32=None, H32=None):
        B = kmat.shape[0]
        n = self.n
        dual = 1 if G32 is not None else 0
        nat.call("mnpbem_green_assemble", dtype, dual, n, nat.ptr(self.pos), nat.ptr(self.nvec),
                 nat.ptr(self.area), nat.ptr(self.mat[side]), nat.ptr(kmat), self.nmat, self.order,
                 self.npairs, nat.ptr(self.prow), nat.ptr(self.pcol), nat.ptr(self.g), nat.ptr(self.f),
                 float(hsign), nat.ptr(G), nat.ptr(H), n, n * n, nat.ptr(G32), nat.ptr(H32), self.gdiv,
                 nat.ptr(self.poff), B, nat.stream_handle(self.device))


class _GroupBEM(BEMRet):
    """BEMRet over the stacked group: factor / solve_batch run unchanged on
    the S L batch; permittivities come per particle."""

    def __init__(self, parts, green, precision, refine_tol, refine_maxit):
        self.parts = parts
        super().__init__(parts[0], precision=precision, green=green, refine_tol=refine_tol,
                         refine_maxit=refine_maxit, residual="stored")

    def _eps_sides(self, enei_list):
        L = self.g.gdiv
        e_in, e_out = np.empty(len(enei_list), complex), np.empty(len(enei_list), complex)
        for s, p in enumerate(self.parts):
            m_in, m_out = p.inout[0, 0], p.inout[0, 1]
            for l in range(L):
                lam = enei_list[s * L + l]
                e_in[s * L + l] = complex(p.eps[m_in - 1](lam)[0])
                e_out[s * L + l] = complex(p.eps[m_out - 1](lam)[0])
        return e_in, e_out


class _Group:
    def __init__(self, idx, parts, greens, precision, refine_tol, refine_maxit):
        self.idx = idx
        self.parts = parts
        self.green = _GroupGreen(greens)
        self.bem = _GroupBEM(parts, self.green, precision, refine_tol, refine_maxit)
        self.n, self.S = self.green.n, self.green.S
        self._flags = None

    def side_flags(self, planewave):
        key = id(planewave)
        if self._flags is None or self._flags[0] != key:
            inm = np.concatenate([(p.inout_faces[:, 0] == planewave.medium) for p in self.parts]).astype(np.int32)
            outm = np.concatenate([(p.inout_faces[:, 1] == planewave.medium) for p in self.parts]).astype(np.int32)
            dev = self.green.device
            self._flags = (key, (torch.as_tensor(inm).to(dev), torch.as_tensor(outm).to(dev),
                                 torch.as_tensor(planewave.pol, dtype=torch.float64).to(dev),
                                 torch.as_tensor(planewave.dir, dtype=torch.float64).to(dev)))
        return self._flags[1]

    def pw_params(self, planewave, lams_b):
        """(S L, 7) {k0, Re k, Im k, eps1, eps2} per batch element."""
        L = self.green.gdiv
        rows = []
        for b, lam in enumerate(lams_b):
            p = self.parts[b // L]
            k0 = 2.0 * np.pi / lam
            k = k0 * np.sqrt(complex(p.eps[planewave.medium - 1](lam)[0]))
            e1 = complex(p.eps[p.inout[0, 0] - 1](lam)[0])
            e2 = complex(p.eps[p.inout[0, 1] - 1](lam)[0])
            rows.append([k0, k.real, k.imag, e1.real, e1.imag, e2.real, e2.imag])
        return torch.as_tensor(np.array(rows), dtype=torch.float64).to(self.green.device)

    def rhs(self, planewave, pw):
        g, P = self.green, planewave.npol
        B, n = pw.shape[0], g.n
        inm, outm, pol, dr = self.side_flags(planewave)
        R = torch.empty((B, n, 8 * P), dtype=torch.complex128, device=g.device)
        nat.call("mnpbem_planewave_batched", n, P, nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(inm), nat.ptr(outm),
                 nat.ptr(pol), nat.ptr(dr), nat.ptr(pw), nat.ptr(R), 8 * P, n * 8 * P, None, 0, g.gdiv, B,
                 nat.stream_handle(g.device))
        return R

    def extinction(self, planewave, X, pw):
        g, P = self.green, planewave.npol
        B, n = pw.shape[0], g.n
        inm, outm, pol, dr = self.side_flags(planewave)
        ext = torch.empty((B, P), dtype=torch.float64, device=g.device)
        nat.call("mnpbem_extinction_batched", n, P, nat.ptr(X), 8 * P, n * 8 * P, nat.ptr(g.pos), nat.ptr(g.area),
                 nat.ptr(inm), nat.ptr(outm), nat.ptr(pol), nat.ptr(dr), nat.ptr(pw), nat.ptr(ext), g.gdiv, B,
                 nat.stream_handle(g.device))
        return ext


class BEMRetBatch:
    """Retarded BEM solver for many ComParticles at once (each with one
    material per side, as BEMRet), batched across the particles of each face
    count.  ``solve_sweep(enei_list, planewave)`` returns the extinction
    spectra, one (n_lambda, npol) array per particle in the given order."""

    def __init__(self, particles, precision="mixed", refine_tol=1e-12, refine_maxit=8, order=5, device="cuda",
                 **options):
        if precision not in ("double", "mixed", "mixed-fp32"):
            raise ValueError("BEMRetBatch: precision must be 'double', 'mixed' or 'mixed-fp32'")
        self.particles = list(particles)
        if not self.particles:
            raise ValueError("BEMRetBatch: no particles")
        self.precision = precision
        self.refine_tol = refine_tol
        self.greens = [CompGreenRet(p, p, order=order, device=device, **options) for p in self.particles]
        self.device = self.greens[0].device
        sizes = np.array([g.n for g in self.greens])
        self.groups = []
        for n in sorted(set(sizes.tolist()), reverse=True):
            idx = np.nonzero(sizes == n)[0]
            self.groups.append(_Group(idx, [self.particles[i] for i in idx], [self.greens[i] for i in idx],
                                      precision, refine_tol, refine_maxit))
        self.last_residual = None

    @property
    def faces(self):
        return [g.n for g in self.greens]

    def block_bytes(self, L):
        """Device bytes of one wavelength block of L for the largest group
        (factors, complex128 copies for the FP64 residual, work space)."""
        es = 16 if self.precision == "double" else 8
        per = lambda n: n * n * (8 * es + (0 if self.precision == "double" else 64))
        return max(g.S * L * per(g.n) for g in self.groups)

    def auto_block(self, nlam, frac=0.5):
        """The largest wavelength block (a divisor of nlam, at most 64) whose
        largest group fits in ``frac`` of the free device memory."""
        free = torch.cuda.mem_get_info(self.device)[0]
        for L in range(min(nlam, 64), 0, -1):
            if nlam % L == 0 and self.block_bytes(L) < frac * free:
                return L
        return 1

    def solve_block(self, lams, planewave, group=None, fixed_iters=None):
        """Extinction (S, L, npol) of one group (or a list for all groups) for
        the wavelength block ``lams``, eager launches."""
        groups = self.groups if group is None else [group]
        out, res = [], []
        for gr in groups:
            gr.green.gdiv = len(lams)
            lb = gr.green.slot_lams(lams)
            F = gr.bem.factor(lb, store=gr.bem.precision != "double")
            pw = gr.pw_params(planewave, lb)
            R = gr.rhs(planewave, pw)
            X = gr.bem.solve_batch(F, R, fixed_iters=fixed_iters)
            out.append(gr.extinction(planewave, X, pw).reshape(gr.S, len(lams), planewave.npol))
            if gr.bem.last_residual is not None and fixed_iters is None:
                res.append(np.max(gr.bem.last_residual))
            del F, X, R
        if res:
            self.last_residual = max(res)
        return out if group is None else out[0]

    def _scatter(self, per_group, nlam, npol):
        spectra = [None] * len(self.particles)
        for gr, e in zip(self.groups, per_group):
            for s, i in enumerate(gr.idx):
                spectra[i] = e[s]
        return spectra

    def solve_sweep(self, enei_list, planewave, block=None):
        """Extinction spectra for every particle: a list of (n_lambda, npol)."""
        lams = np.atleast_1d(np.asarray(enei_list, dtype=float))
        L = block or self.auto_block(len(lams))
        parts = [[] for _ in self.groups]
        for c0 in range(0, len(lams), L):
            for k, e in enumerate(self.solve_block(lams[c0:c0 + L], planewave)):
                parts[k].append(e.cpu().numpy())
        per_group = [np.concatenate(p, axis=1) for p in parts]
        return self._scatter(per_group, len(lams), planewave.npol)


class BatchSweepGraph:
    """One wavelength block of every group of a BEMRetBatch (assembly,
    factorisations, right sides, ``iters`` FP64 refinement steps, extinction)
    captured once as a HIP graph and replayed for any block of ``L``
    wavelengths.  Inputs (wavenumbers, permittivities, plane-wave parameters
    per batch element) are copied into device buffers before each replay; the
    captured region transfers nothing from the host, allocates only from the
    graph's pool and queries nothing.  ``run`` reads the final residual norms
    after the replay and raises if any (particle, wavelength) misses the
    solver's refine_tol."""

    def __init__(self, batch, planewave, L, iters=3):
        if batch.precision == "double":
            raise ValueError("BatchSweepGraph: the captured step is the mixed-precision path")
        self.batch, self.pw, self.L, self.iters = batch, planewave, int(L), int(iters)
        dev = batch.device
        self.inputs = []
        for gr in batch.groups:
            B = gr.S * self.L
            self.inputs.append({"kmat": torch.empty((B, gr.green.nmat, 2), dtype=torch.float64, device=dev),
                                "par": torch.empty((B, 5), dtype=torch.float64, device=dev),
                                "pw": torch.empty((B, 7), dtype=torch.float64, device=dev)})
        self.graph = None
        self.ext, self.norms = None, None
        self.fallbacks = 0  # groups of a replayed block solved again eagerly

    def _fill(self, lams):
        for gr, buf in zip(self.batch.groups, self.inputs):
            gr.green.gdiv = self.L
            lb = gr.green.slot_lams(lams)
            e_in, e_out = gr.bem._eps_sides(lb)
            buf["kmat"].copy_(gr.green.wavenumbers(lb))
            buf["par"].copy_(_par_tensor(e_in, e_out, 2.0 * np.pi / lb, self.batch.device))
            buf["pw"].copy_(gr.pw_params(self.pw, lb))

    def _step(self, lams):
        ext, norms, keep, flags = [], [], [], []
        for gr, buf in zip(self.batch.groups, self.inputs):
            gr.green.gdiv = self.L
            keep.append(gr.side_flags(self.pw))
            lb = gr.green.slot_lams(lams)
            F = gr.bem.factor(lb, kmat=buf["kmat"], par=buf["par"], store=True)
            R = gr.rhs(self.pw, buf["pw"])
            X = gr.bem.solve_batch(F, R, fixed_iters=self.iters)
            ext.append(gr.extinction(self.pw, X, buf["pw"]))
            norms.append(gr.bem.last_norms)
            flags.append(F.npflag)
            del F, R, X
        self.ext, self.norms, self._keep, self.npflags = ext, norms, keep, flags

    def run(self, lams):
        """Extinction (S, L, npol) per group for the block ``lams``."""
        lams = np.atleast_1d(np.asarray(lams, dtype=float))
        if len(lams) != self.L:
            raise ValueError("BatchSweepGraph: captured for %d wavelengths, got %d" % (self.L, len(lams)))
        self._fill(lams)
        if self.graph is None:
            dev = self.batch.device
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                self._step(lams)   # eager warm-up on a side stream (library, caches)
            torch.cuda.current_stream(dev).wait_stream(side)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._step(lams)
        self.graph.replay()
        groups = self.batch.groups
        out = [e.reshape(gr.S, self.L, self.pw.npol) for gr, e in zip(groups, self.ext)]
        worst, redo = 0.0, []
        for gi, (nrm, fl) in enumerate(zip(self.norms, self.npflags)):
            if needs_interchanges(fl):
                # the captured factorisation ran without the row interchanges
                # that this group's block needs: the eager path factors it
                # with them
                redo.append(gi)
                continue
            h = nrm.cpu().numpy()
            w = float(np.max(np.sqrt(h[:, 0] / np.maximum(h[:, 1], 1e-300))))
            if w > self.batch.refine_tol:
                if fl is None:
                    raise RuntimeError("BatchSweepGraph: residual %.2e above refine_tol after %d steps; use more "
                                       "iters" % (w, self.iters))
                # Sigma without interchanges needs more steps than captured:
                # this group alone is solved again eagerly (its refinement
                # runs until refine_tol), the other groups keep the replay's
                redo.append(gi)
                continue
            worst = max(worst, w)
        for gi in redo:
            out[gi] = self.batch.solve_block(lams, self.pw, group=groups[gi])
            worst = max(worst, self.batch.last_residual)
        self.fallbacks += len(redo)
        self.last_residual = worst
        return out

    def solve_sweep(self, enei_list):
        """Extinction spectra of every particle over ``enei_list`` (a multiple
        of L wavelengths), one replay per block."""
        lams = np.atleast_1d(np.asarray(enei_list, dtype=float))
        if len(lams) % self.L:
            raise ValueError("BatchSweepGraph: %d wavelengths is not a multiple of the block %d" % (len(lams), self.L))
        parts = [[] for _ in self.batch.groups]
        for c0 in range(0, len(lams), self.L):
            for k, e in enumerate(self.run(lams[c0:c0 + self.L])):
                parts[k].append(e.cpu().numpy())
        per_group = [np.concatenate(p, axis=1) for p in parts]
        return self.batch._scatter(per_group, len(lams), self.pw.npol)
