"""The LDS FFT design of sar_cambrianaccelerationcard_amd/csrc/sarbp_fft.h,
emulated on the CPU: the Stockham pass plans (radix 8 on n/8 threads, radix
16 on n/16 threads, one smaller pass first) and their twiddle indexing give
the DFT, and the LDS slot maps (one pad word per 16 up to 4096 points, an XOR
swizzle at 8192 and 16384) keep each 16-lane access of every pass within at most two
lanes per bank pair.  The device kernels themselves are checked on the GPU
through range compression and omega-K parity.  The register-resident form
(fft_regs: 8 or 16 points per thread, the first and last pass on registers)
is emulated with its thread ownership of points and checked the same way."""
import numpy as np
import pytest


def plan(lg, radix):
    k = {8: 3, 16: 4}[radix]
    first = lg % k
    return ([1 << first] if first else []) + [radix] * (lg // k)


def slot(i, lg):
    return i + (i >> 4) if lg <= 12 else i ^ (((i >> 4) & 7) << 1)


def stockham(a, radices, inv=False):
    n = a.size
    lg = n.bit_length() - 1
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    buf = np.zeros(n + n // 16 + 1, complex)
    for i in range(n):
        buf[slot(i, lg)] = a[i]
    lns = 0
    for R in radices:
        lgr = R.bit_length() - 1
        nb, ns = n // R, 1 << lns
        W = np.exp((2j if inv else -2j) * np.pi * np.outer(np.arange(R), np.arange(R)) / R)
        b = np.arange(nb)
        k = b & (ns - 1)
        x = np.stack([buf[[slot(int(i), lg) for i in b + r * nb]] for r in range(R)])
        if lns:
            w1 = tw[k << (lg - lns - lgr)]
            x = x * (np.conj(w1) if inv else w1)[None, :] ** np.arange(R)[:, None]
        y = W @ x
        dst = ((b - k) << lgr) + k
        for r in range(R):
            buf[[slot(int(i), lg) for i in dst + r * ns]] = y[r]
        lns += lgr
    return np.array([buf[slot(i, lg)] for i in range(n)])


@pytest.mark.parametrize("lg,radix", [(lg, r) for lg in range(1, 14) for r in (8, 16)] + [(14, 16)])
def test_plans_compute_the_dft(lg, radix):
    r = np.random.default_rng(lg)
    a = r.standard_normal(1 << lg) + 1j * r.standard_normal(1 << lg)
    for inv in (False, True):
        got = stockham(a, plan(lg, radix), inv)
        ref = np.fft.ifft(a) * a.size if inv else np.fft.fft(a)
        assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def _worst(addrs):
    # 16 lanes, 32 banks of 4 B: a complex64 word at slot s uses banks 2s, 2s+1.
    per = {}
    for s in addrs:
        per.setdefault(s % 16, set()).add(s)
    return max(len(v) for v in per.values())


def _pass_conflicts(lg, radix, mapf):
    n = 1 << lg
    lns, worst = 0, []
    for R in plan(lg, radix):
        lgr = R.bit_length() - 1
        nb, ns = n // R, 1 << lns
        w = rd = 1
        for g in range(0, nb, 16):
            b = np.arange(g, g + 16)
            k = b & (ns - 1)
            for r in range(R):
                w = max(w, _worst([mapf(int(i)) for i in ((b - k) << lgr) + k + r * ns]))
                rd = max(rd, _worst([mapf(int(i)) for i in b + r * nb]))
        worst.append((R, ns, w, rd))
        lns += lgr
    return worst


@pytest.mark.parametrize("lg,radix", [(12, 8), (13, 16), (13, 8), (14, 16)])
def test_slot_maps_spread_the_stockham_strides(lg, radix):
    mapped = _pass_conflicts(lg, radix, lambda i: slot(i, lg))
    assert all(w <= 2 and rd <= 2 for _, _, w, rd in mapped), mapped
    # Unmapped, some pass writes 4 lanes or more into one bank pair.
    plain = _pass_conflicts(lg, radix, lambda i: i)
    assert max(w for _, _, w, _ in plain) >= 4, plain


def _dft(R, inv):
    return np.exp((2j if inv else -2j) * np.pi * np.outer(np.arange(R), np.arange(R)) / R)


def _pass(v, R, lns, inv):
    """One Stockham pass in natural indexing (stockham_pass without slots)."""
    n = v.size
    lg, lgr = n.bit_length() - 1, R.bit_length() - 1
    nb, ns = n // R, 1 << lns
    b = np.arange(nb)
    k = b & (ns - 1)
    x = np.stack([v[b + r * nb] for r in range(R)])
    if lns:
        w1 = np.exp(-2j * np.pi * (k << (lg - lns - lgr)) / n)
        x = x * (np.conj(w1) if inv else w1)[None, :] ** np.arange(R)[:, None]
    y = _dft(R, inv) @ x
    out = np.empty_like(v)
    dst = ((b - k) << lgr) + k
    for r in range(R):
        out[dst + r * ns] = y[r]
    return out


def fft_regs(a, lg, pt, inv, transposed_first=False):
    """fft::fft_regs<lg, inv, pt>: thread t holds x[t][e] = point t + e T
    (T = n / pt) on entry and output t + e T on return.  The first pass's
    butterfly p of thread t takes x[t][p + r pt/R0] (transposed_first plants
    the mistake x[t][p R0 + r]); the last radix-pt pass's butterfly t reads
    t + r T and keeps output r in x[t][r]."""
    n = 1 << lg
    T = n // pt
    radices = plan(lg, pt)
    R0 = radices[0]
    per0, lgr0 = pt // R0, R0.bit_length() - 1
    x = a.reshape(pt, T).T                       # x[t][e] = a[t + e T]
    buf = np.empty(n, complex)
    for p in range(per0):
        cols = [p * R0 + r if transposed_first else p + r * per0 for r in range(R0)]
        y = _dft(R0, inv) @ np.stack([x[:, c] for c in cols])
        b = np.arange(T) + p * T
        for r in range(R0):
            buf[(b << lgr0) + r] = y[r]
    lns = lgr0
    for R in radices[1:-1]:
        buf = _pass(buf, R, lns, inv)
        lns += R.bit_length() - 1
    assert lns == lg - (pt.bit_length() - 1)
    t = np.arange(T)
    xs = np.stack([buf[t + r * T] for r in range(pt)])
    w1 = np.exp(-2j * np.pi * t / n)             # k = t, tw[k << 0]
    xs = xs * (np.conj(w1) if inv else w1)[None, :] ** np.arange(pt)[:, None]
    y = _dft(pt, inv) @ xs                        # y[r][t] = output t + r T
    return y.reshape(n)


@pytest.mark.parametrize("lg,pt", [(lg, 16) for lg in range(8, 15)] + [(lg, 8) for lg in range(6, 14)])
def test_register_fft_computes_the_dft(lg, pt):
    r = np.random.default_rng(100 + lg)
    a = r.standard_normal(1 << lg) + 1j * r.standard_normal(1 << lg)
    for inv in (False, True):
        got = fft_regs(a, lg, pt, inv)
        ref = np.fft.ifft(a) * a.size if inv else np.fft.fft(a)
        assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("lg,pt", [(13, 16), (14, 16), (13, 8), (11, 8)])  # several first-pass butterflies per thread
def test_register_fft_catches_a_transposed_first_pass(lg, pt):
    r = np.random.default_rng(7)
    a = r.standard_normal(1 << lg) + 1j * r.standard_normal(1 << lg)
    got = fft_regs(a, lg, pt, False, transposed_first=True)
    assert np.max(np.abs(got - np.fft.fft(a))) > 1e-3 * np.max(np.abs(a)) * (1 << lg) ** 0.5
