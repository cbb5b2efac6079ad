"""CPU oracle for the MHDFlows.jl incompressible HD/MHD right-hand side and time steppers.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library under
``mhdflows.jl_amd/csrc`` and the host package ``mhdflows_amd``) imports, links or
calls this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and there only as the checker
(or as the timed CPU restatement in ``cpu_baseline``), never as the thing measured.

What it is
----------
A numpy + scipy.fft (pocketfft) restatement of the reference path, in Float32 and
Float64, written from the reference's Julia sources (MHDFlows.jl v0.2.2):

* ``grid``   -- FourierFlows ``ThreeDGrid`` conventions as mirrored by
  ``src/utils/utils.jl:42-98`` (wavenumbers, ``invKrsq[1,1,1]=0``, alias ranges).
* ``rhs``    -- ``MHDcalcN!`` (``src/pgen.jl:153-162``), ``MHDcalcN_advection!``
  (``src/Solver/MHDSolver.jl:330-351``), ``UᵢUpdate!`` (``MHDSolver.jl:27-103``),
  ``BᵢUpdate!`` (``MHDSolver.jl:106-177``), and the HD path ``HDcalcN!``
  (``src/pgen.jl:173-181``) / ``HDSolver.jl:25-108``.  Two modes: ``literal`` keeps
  the reference's operation structure (36 FFTs per MHD right-hand side, one
  product at a time, per-direction accumulation); ``fused`` evaluates the same
  mathematics with 15 FFTs.
* ``steppers`` -- RK4 as restated in-repo at ``src/DyeModule.jl:62-91`` and
  LSRK54 (Carpenter & Kennedy 1994, five-stage 2N-storage, 4th order).
* ``diagnostics`` -- ``getCFL!`` (``src/integrator.jl:158-198``), ``∑²`` /
  ``ProbDiagnostic`` (``src/utils/UserInterface.jl:29,65-86``), ``Div``
  (``src/utils/VectorCalculus.jl:72-94``), helicities (``src/utils/MHDAnalysis.jl:94-174``).
* ``integrator`` -- the ``TimeIntegrator!`` hot loop (``src/integrator.jl:31-156``).

How each part is pinned (checked by ``tests/test_oracle_*.py``)
----------------------------------------------------------------
* R2C/C2R conventions: brute-force O(N^2) DFT at 4^3..8^3 and the Hermitian
  projection identity ``rfftn(irfftn(X)) == P_H(X)``.
* Projection / Lorentz / induction / diffusion algebra: closed forms -- single-mode
  viscous/resistive decay (RK4 factor ``1+z+z^2/2+z^3/6+z^4/24`` per step), the
  linear Alfven wave on a uniform field along each axis (exact RK4 matrix
  polynomial), divergence of the right-hand side at round-off, ideal-MHD energy
  and cross-helicity conservation to O(dt^4).
* Literal vs fused modes agree to round-off (two independent op structures).
* HD sign: the HD right-hand side equals the MHD velocity right-hand side with b = 0
  (whose u u sign cross-helicity conservation fixes), and the 3D Taylor-Green field's
  t = 0 right-hand side matches its closed form (pressure (cos 2x + cos 2y)(2 + cos 2z)/16).
* Helicities: Arnold-Beltrami-Childress fields (curl f = k f) give H_kin = k E_kin and
  H_mag = E_mag / k, which fixes the signs of Curl and VectorPotential and the dV factor.
* Forcing modules (``forcing``): N97 against its closed form and support; negative damping
  against its injected power (sum u.F dV = P for f >= 0); A99 against its construction
  (solenoidal, orthonormal e1/e2, zero kx = 0 plane, amplitude ~ sqrt(P)) and the uniformity
  and independence of its counter-based draws.  A99's random stream itself is parity unpinned:
  the reference draws unseeded rand(), so the build defines the draws (oracle and device share
  them) and only the distribution and construction can be compared.
* Structure functions (``analysis.cf``, ``sf2_1d``, ``sfr_global``, TurbStatTool.jl): CF against a
  brute-force correlation sum, the binning against the same sums taken over signed lags with the
  reference's idx = lag + 1, and the global binning against an axis-aligned mean field.
* Divergence corrections (``vp``): the Helmholtz split (gradient removed, solenoidal part kept,
  k . F = 0 after) and idempotence.
* LSRK54: order conditions of the stability polynomial (Taylor series through z^4).
* Alias index table: the kept ranges listed for the FourierFlows 1/3 rule.

Parity unpinned (no import of the reference is possible here: Julia,
FourierFlows.jl, FFTW.jl and CUDA.jl are absent, and the reference has no tests
or golden vectors):

1. FourierFlows ``getaliasedwavenumbers`` index endpoints (recalled v0.10.1 formula,
   pinned only by the kept-range table it reproduces).
2. FourierFlows RK4 / LSRK54 operation order and stage times (RK4 follows the
   in-repo restatement ``DyeModule.jl:62-91``; LSRK54 coefficients pinned by order
   conditions only).
3. FFTW's rounding and its non-Hermitian c2r behaviour (pocketfft is used; the
   Hermitian-projection semantics are pinned mathematically, not against FFTW).
4. Julia broadcast promotion order in Float32 runs (``ν``, ``η`` kept Float64 as
   ``MHDParams.jl:45-47`` types them ``Number``).
5. The ``ThreeDGrid`` default origin ``x0 = -L/2`` (affects initial-condition
   sampling only).
"""

from . import grid, rhs, steppers, diagnostics, integrator  # noqa: F401
from .problem import OracleProblem  # noqa: F401
