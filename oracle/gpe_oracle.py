"""FP64 NumPy oracle of the reference split-step path (test infrastructure).

THIS MODULE IS A CHECKER, NOT PART OF THE PRODUCT.  Only tests/, the smoke
check in __graft_entry__.py and bench.py's cpu_baseline leg import it.  The
engine (gpe_simulator_amd) never imports, calls or falls back to it.

It restates, in complex128 / float64 with numpy.fft (pocketfft), the
algorithm of PoyntingZhang/GPE_simulator's propagation path.  Each function
cites the reference file:line it follows.  Equation (reference units):
    i dpsi/dt = [-lap + V + beta |psi|^2] psi     (no 1/2 on the Laplacian)

Pins (tests/test_oracle.py): the reference's own closed-form tests
(tests/test_propagators.py:61-158, tests/test_diagnostics.py:31-81), exact
time reversibility, Parseval, a brute-force DFT at N <= 6, the harmonic
ground-state energy E = mu = 3, plane-wave mu and E, the integer shell
identity and the stage clock t = i*dt (tests/test_workflow.py:262).  The
equal-width spectrum view (no reference counterpart) is pinned against a
50-digit decimal evaluation of its defining formula.  The Gaussian start
against the continuum normalisation pi^(-3/4) sigma^(-3/2) and its profile,
beta against the workload table (beta = 200, 1000, 1100, 4000 at the listed
scattering lengths), every expression function against Python's math, the
box and absorber volumes against hand-evaluated points (wall maxima, face
ramps, corner sums), and the monitor images against direct sums, Parseval
and symmetry.
"""

from __future__ import annotations

import ast
import math

import numpy as np

A0 = 5.29177210544e-11
LENGTH_UNIT = 1e-6

# FFT backend: numpy.fft (single-threaded pocketfft) by default; tests may
# switch to scipy.fft with worker threads for speed (same algorithm family,
# FP64 throughout).
_fftn, _ifftn = np.fft.fftn, np.fft.ifftn


def use_threads(workers):
    """Run the oracle's 3D FFTs on `workers` threads (None: numpy, 1 thread)."""
    global _fftn, _ifftn
    if workers is None:
        _fftn, _ifftn = np.fft.fftn, np.fft.ifftn
        return
    import scipy.fft
    _fftn = lambda a: scipy.fft.fftn(a, workers=workers)  # noqa: E731
    _ifftn = lambda a: scipy.fft.ifftn(a, workers=workers)  # noqa: E731


def get_grids(n, length):
    """x = linspace(-L/2, L/2, N, endpoint=False); k = 2*pi*fftfreq(N, L/N).
    gpe/grids.py:3-6."""
    x = np.linspace(-length / 2, length / 2, n, endpoint=False)
    k = 2 * np.pi * np.fft.fftfreq(n, length / n)
    return x, k


def get_beta(n_atoms, a):
    """beta = 8*pi*N_atom*a*a0/l.  gpe/helper.py:109-113."""
    return 8 * np.pi * n_atoms * (a * A0) / LENGTH_UNIT


def get_N(psi, x):
    """dx*dx*dx * sum |psi|^2.  gpe/helper.py:51-88 (sum_partial/sum_finish)."""
    dx = x[1] - x[0]
    return dx * dx * dx * float(np.sum(psi.real * psi.real + psi.imag * psi.imag))


def normalise(psi, x):
    """psi /= sqrt(N).  gpe/helper.py:91-93."""
    psi /= np.sqrt(get_N(psi, x))
    return psi


def gaussian(x, sigma):
    """Separable exp(-x^2/(2 sigma^2)) product, normalised.
    gpe/initial_states.py:4-14."""
    g = np.exp(-x ** 2 / (2 * sigma ** 2)).astype(np.complex128)
    psi = g[:, None, None] * g[None, :, None] * g[None, None, :]
    return normalise(psi, x)


# ------------------------------------------------------------ expressions --
_FUNCS = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "sqrt": np.sqrt,
          "tanh": np.tanh, "abs": np.abs, "log": np.log}


def eval_expression(source, x, t=0.0):
    """Evaluate a schema-2 potential expression on the grid (x, y, z from the
    same axis), in float64, as the spliced C of gpe/make_kernel.py:14-45 and
    gpe/expressions.py:9-30 computes it pointwise."""
    tree = ast.parse(str(source).strip(), mode="eval").body
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]

    def ev(node):
        if isinstance(node, ast.Constant):
            return float(node.value)
        if isinstance(node, ast.Name):
            return {"x": X, "y": Y, "z": Z, "t": float(t), "pi": math.pi}[node.id]
        if isinstance(node, ast.UnaryOp):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.Call):
            return _FUNCS[node.func.id](ev(node.args[0]))
        a, b = ev(node.left), ev(node.right)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):
            return a / b
        return np.power(a, b)

    return np.broadcast_to(ev(tree), (len(x),) * 3).astype(np.float64)


# ----------------------------------------------------- static volumes --
def box_potential(x, box_radius, box_length, wall_thickness, trap_depth):
    """Cylindrical box walls: depth * max(tube, sheet) with
    tube = exp(-(sqrt(x^2+y^2) - R)^2 / (2 w^2)) and
    sheet = exp(-(|z| - L/2)^2 / (2 w^2)).  gpe/potentials.py:3-12."""
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    rho = np.sqrt(X * X + Y * Y)
    tube = np.exp(-(rho - box_radius) ** 2 / (2 * wall_thickness ** 2))
    sheet = np.exp(-(np.abs(Z) - box_length / 2) ** 2 / (2 * wall_thickness ** 2))
    return trap_depth * np.maximum(tube, sheet)


def absorption_potential(activation_distance, amplitude, x, simulation_range):
    """Linear ramps inside `activation_distance` of each of the six faces of
    the box [-range/2, range/2)^3, summed and scaled by `amplitude`.
    gpe/potentials.py:14-28."""
    def ramps(c):
        return (np.maximum(activation_distance - (c + simulation_range / 2), 0)
                + np.maximum(activation_distance - (simulation_range / 2 - c), 0))
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    return amplitude * ((ramps(X) + ramps(Y)) + ramps(Z))


def total_static_potential(real_grid, imag_grid):
    """Complex static volume real - i*imag (the absorber enters with a
    negative imaginary part).  gpe/potentials.py:31-37."""
    return np.asarray(real_grid, dtype=np.float64) - 1j * np.asarray(imag_grid, dtype=np.float64)


# --------------------------------------------------------- monitor images --
def monitor_images(psi, x, k):
    """The monitor arrays of gpe/save.py:59-90 from the completed psi:
    TOF projections sum_abs2(fftn(psi)) over z and over x times dx^3 dk /
    (size dk^3), the kz = 0 and kx = 0 momentum slices times dx^3 / (size
    dk^3), all fftshifted; the real-space column density sum_z |psi|^2 dx;
    the mid-z plane's density and phase (NaN where the density is below 1e-12
    of its maximum).  sum_abs2 is gpe/helper.py:3-12."""
    n = psi.shape[0]
    dx = float(x[1] - x[0])
    dk = abs(float(k[1] - k[0]))
    psi_k = _fftn(psi)
    power = psi_k.real ** 2 + psi_k.imag ** 2
    scale = dx ** 3 / (psi.size * dk ** 3)
    shift = np.fft.fftshift
    plane = psi[:, :, n // 2]
    density = plane.real ** 2 + plane.imag ** 2
    return dict(
        TOF_xy=shift(power.sum(axis=2)) * (scale * dk),
        TOF_yz=shift(power.sum(axis=0)) * (scale * dk),
        TOF_slice_xy=shift(power[:, :, 0]) * scale,
        TOF_slice_yz=shift(power[0, :, :]) * scale,
        density_xy=(psi.real ** 2 + psi.imag ** 2).sum(axis=2) * dx,
        density_slice_xy=density,
        phase_xy=np.where(density > density.max() * 1e-12, np.angle(plane), np.nan))


# ------------------------------------------------------------ local steps --
def real_half_step(psi, v_static, potential, beta, dt):
    """psi *= exp(Im V dt/2) * exp(-i (Re V + beta|psi|^2 + U + W) dt/2).
    gpe/make_kernel.py:14-45 (dynamic) and :54-83 (static)."""
    abs2 = psi.real * psi.real + psi.imag * psi.imag
    real_phase = (v_static.real + beta * abs2 + potential) * (dt / 2)
    imag_phase = v_static.imag * (dt / 2)
    psi *= np.exp(imag_phase - 1j * real_phase)
    return psi


def imag_half_step(psi, v_static, potential, beta, dt):
    """psi *= exp(-(V + beta|psi|^2 + U) dtau/2).  gpe/make_kernel.py:93-119."""
    abs2 = psi.real * psi.real + psi.imag * psi.imag
    psi *= np.exp(-(v_static + beta * abs2 + potential) * (dt / 2))
    return psi


def k_squared(k):
    return k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2


def kinetic(psi_k, k, dt, kcut=None):
    """psi_k *= exp(-i k^2 dt); with a cutoff, modes k^2 > kcut^2 are zeroed.
    gpe/kernels.py:3-21, :24-45."""
    k2 = k_squared(k)
    psi_k *= np.exp(-1j * k2 * dt)
    if kcut is not None:
        psi_k[k2 > kcut * kcut] = 0
    return psi_k


def imag_kinetic(psi_k, k, dt):
    """psi_k *= exp(-k^2 dtau).  gpe/kernels.py:48-66."""
    psi_k *= np.exp(-k_squared(k) * dt)
    return psi_k


class Problem:
    """Everything one propagator closure reads from a stage
    (gpe/make_propagator.py:12-19, 57-68, 109-116)."""

    def __init__(self, n, length, kind, dt, beta, static=None, dynamic=None, kcut=None,
                 v_static=None):
        self.n, self.length, self.kind, self.dt, self.beta = n, length, kind, dt, beta
        self.x, self.k = get_grids(n, length)
        self.static_source, self.dynamic_source, self.kcut = static, dynamic, kcut
        self.U = eval_expression(static, self.x) if static else np.zeros((n,) * 3)
        if v_static is None:
            v_static = np.zeros((n,) * 3, dtype=np.complex128 if kind == "real" else np.float64)
        self.v_static = v_static

    def W(self, t):
        if not self.dynamic_source:
            return 0.0
        return eval_expression(self.dynamic_source, self.x, t)

    def step(self, psi, t=None):
        """One Strang step; returns the new psi.  gpe/make_propagator.py:48-53,
        :94-102 (t_mid = t + dt/2 for both halves) and :143-150 (normalise)."""
        dt = self.dt
        if self.kind == "real":
            potential = self.U
            if self.dynamic_source:
                t_mid = float(t) + dt / 2
                potential = potential + self.W(t_mid)
            psi = real_half_step(psi, self.v_static, potential, self.beta, dt)
            psi_k = kinetic(_fftn(psi), self.k, dt, self.kcut)
            psi = _ifftn(psi_k)
            return real_half_step(psi, self.v_static, potential, self.beta, dt)
        psi = imag_half_step(psi, self.v_static, self.U, self.beta, dt)
        psi_k = imag_kinetic(_fftn(psi), self.k, dt)
        psi = _ifftn(psi_k)
        psi = imag_half_step(psi, self.v_static, self.U, self.beta, dt)
        return normalise(psi, self.x)

    def run(self, psi, steps, first=0):
        """launch_stage's loop, t_i = i*dt.  gpe/launch.py:4-14."""
        psi = np.array(psi, dtype=np.complex128, copy=True)
        for i in range(first, first + steps):
            psi = self.step(psi, i * self.dt if self.dynamic_source else None)
        return psi

    # -- diagnostics
    def energies(self, psi, t=0.0):
        """norm, E_kin (gpe/helper.py:135-140), E_pot, E_int, E, mu, parseval.

        E = int(|grad psi|^2 + V|psi|^2 + beta/2 |psi|^4),
        mu = int(|grad psi|^2 + V|psi|^2 + beta |psi|^4) / int |psi|^2 —
        quantities the reference does not define; the kinetic term is the
        reference's E_kin."""
        x = self.x
        dx = x[1] - x[0]
        dx3 = dx * dx * dx
        a2 = psi.real ** 2 + psi.imag ** 2
        potential = self.U + (self.v_static.real if np.iscomplexobj(self.v_static) else self.v_static)
        if self.dynamic_source:
            potential = potential + self.W(t)
        F = _fftn(psi)
        f2 = F.real ** 2 + F.imag ** 2
        norm = dx3 * float(np.sum(a2))
        ekin = float(np.sum(f2 * k_squared(self.k))) * dx ** 3 / psi.size
        epot = dx3 * float(np.sum(potential * a2))
        eint = 0.5 * self.beta * dx3 * float(np.sum(a2 * a2))
        return dict(norm=norm, E_kin=ekin, E_pot=epot, E_int=eint, E=ekin + epot + eint,
                    mu=(ekin + epot + 2 * eint) / norm, parseval=dx3 * float(np.sum(f2)) / psi.size)


def get_E(psi_k, x, k):
    """E_kin = sum |F|^2 (kx^2+ky^2+kz^2) * dx^3 / size.  gpe/helper.py:135-140."""
    dx = x[1] - x[0]
    return float(np.sum((psi_k.real ** 2 + psi_k.imag ** 2) * k_squared(k))) * dx ** 3 / psi_k.size


def nk_extraction(psi_k, x, k):
    """Shell histogram of raw fftn(psi).  gpe/nk_extraction.py:5-45:
    shell = floor(sqrt(k^2)/dk + 0.5), nbins = floor(sqrt(3)*(N//2)+0.5)+1."""
    n = psi_k.shape[0]
    dx = float(x[1] - x[0])
    dk = abs(float(k[1] - k[0]))
    nbins = int(math.floor(math.sqrt(3) * (n // 2) + 0.5)) + 1
    shell = np.floor(np.sqrt(k_squared(k)) / dk + 0.5).astype(np.int64).ravel()
    occupation = (psi_k.real ** 2 + psi_k.imag ** 2).ravel()
    counts = np.bincount(shell, minlength=nbins)
    occ = np.bincount(shell, weights=occupation, minlength=nbins)
    Nk = occ * (dx ** 3 / psi_k.size)
    nk = Nk / (np.maximum(counts, 1) * dk ** 3)
    kk = np.arange(nbins, dtype=np.float64) * dk
    edges = (np.arange(nbins + 1, dtype=np.float64) - 0.5) * dk
    edges[0] = 0
    complete = edges[1:] <= n * dk / 2
    return dict(k=kk, nk=nk, Nk=Nk, counts=counts, edges=edges, complete_shells=complete)


def shell_index_integer(p, q, r):
    """Integer form of the shell index: m = isqrt(s), s = p^2+q^2+r^2, and
    floor(sqrt(s) + 1/2) = m + (s > m^2 + m)."""
    s = np.asarray(p) ** 2 + np.asarray(q) ** 2 + np.asarray(r) ** 2
    m = np.floor(np.sqrt(s.astype(np.float64))).astype(np.int64)
    m = np.where(m * m > s, m - 1, m)
    m = np.where((m + 1) * (m + 1) <= s, m + 1, m)
    return m + (s > m * m + m)


VIEW_BINS = 128


def _isqrt(x):
    """floor(sqrt(x)) of a non-negative int64 array, exact."""
    m = np.floor(np.sqrt(x.astype(np.float64))).astype(np.int64)
    m = np.where(m * m > x, m - 1, m)
    return np.where((m + 1) * (m + 1) <= x, m + 1, m)


def spectrum_view_bin(p, q, r, n, nview=VIEW_BINS):
    """Bin of the equal-width spectrum view (no reference counterpart; the
    reference bins by dk shells, gpe/nk_extraction.py:16-45): with integer
    frequencies m = (p, q, r) and h = n // 2, bin floor(nview |m| / (sqrt(3) h)),
    i.e. |k| / k_max in nview bins, k_max = sqrt(3) h dk the corner; the corner
    itself goes to the last bin.  Evaluated as isqrt(floor(X)) with
    X = nview^2 |m|^2 / (3 h^2), since floor(sqrt(X)) = isqrt(floor(X))."""
    s = (np.asarray(p, dtype=np.int64) ** 2 + np.asarray(q, dtype=np.int64) ** 2
         + np.asarray(r, dtype=np.int64) ** 2)
    return view_bin_from_square(s, n, nview)


def view_bin_from_square(s, n, nview=VIEW_BINS):
    """spectrum_view_bin as a function of the integer |m|^2 = s."""
    h = n // 2
    b = _isqrt((nview * nview * np.asarray(s, dtype=np.int64)) // (3 * h * h))
    return np.minimum(b, nview - 1)


def spectrum_view(psi_k, x, k, nview=VIEW_BINS):
    """|F|^2 of raw fftn(psi) in the equal-width view: counts, raw sums and the
    shell-style normalisations Nk = sum dx^3/N^3, nk = Nk / (count dk^3)."""
    n = psi_k.shape[0]
    dx = float(x[1] - x[0])
    dk = abs(float(k[1] - k[0]))
    m = np.fft.fftfreq(n, 1.0 / n).round().astype(np.int64)
    b = spectrum_view_bin(m[:, None, None], m[None, :, None], m[None, None, :], n, nview).ravel()
    f2 = (psi_k.real ** 2 + psi_k.imag ** 2).ravel()
    counts = np.bincount(b, minlength=nview)
    occupation = np.bincount(b, weights=f2, minlength=nview)
    Nk = occupation * (dx ** 3 / psi_k.size)
    k_max = math.sqrt(3) * (n // 2) * dk
    width = k_max / nview
    return dict(counts=counts, occupation=occupation, Nk=Nk, nk=Nk / (np.maximum(counts, 1) * dk ** 3),
                k=(np.arange(nview) + 0.5) * width, edges=np.arange(nview + 1) * width, k_max=k_max)


def synthetic_state(n, length, seed=0, beta=1.0, potential=None):
    """Seeded Thomas-Fermi-like state plus small complex noise, normalised."""
    x, _ = get_grids(n, length)
    rng = np.random.default_rng(seed)
    if potential is None:
        potential = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    mu = np.percentile(potential, 5)
    psi = np.sqrt(np.maximum(mu - potential, 0) / max(beta, 1e-12)).astype(np.complex128)
    psi += 1e-3 * (rng.standard_normal((n,) * 3) + 1j * rng.standard_normal((n,) * 3))
    return normalise(psi, x)
