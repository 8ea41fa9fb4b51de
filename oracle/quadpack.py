"""Test infrastructure: a restatement of the adaptive quadrature the reference calls for the 1-D
integrals of its shape transforms, so that the library's converged integrals can be compared with
what the reference computes.

The reference evaluates every integral of characteristic_function_fourier_integrals.c with
gsl_integration_qag(f, lo, hi, EPS_ABS, EPS_REL, LIMIT, key = 6) (e.g. :130, :241 and :246, the real
and imaginary parts of a complex integrand separately), with EPS_ABS = EPS_REL = 1e-4 and LIMIT = 1000
(8_band.h:61-63).  GSL's qag is QUADPACK's QAG (Piessens, de Doncker-Kapenga, Ueberhuber and Kahaner,
QUADPACK, Springer 1983, routines DQAGE and DQK61; QUADPACK is public domain), restated here from
that published description:

  * the 61-point Gauss-Kronrod rule (the 30-point Gauss rule and its Kronrod extension), its error
    estimate |K61 - G30| rescaled as resasc min(1, (200 err / resasc)^1.5) and floored at
    50 eps resabs;
  * globally adaptive bisection of the interval with the largest error estimate until the summed
    estimate is at most max(epsabs, epsrel |result|), with QUADPACK's round-off and limit stops.

The Kronrod nodes and weights are computed here from their definition, not tabulated (Kronrod
1965: the added nodes are the zeros of the Stieltjes polynomial; see gauss_kronrod).
tests/test_quadpack_cpu.py pins the rule (it integrates every polynomial of degree <= 3N+1 = 91
exactly although only degrees <= 2N fix its weights, it is not the 61-point Gauss rule, positive
weights, interlacing nodes) and the adaptive driver (closed-form integrals, forced subdivision).
"""
import math

import numpy as np

EPS_ABS = 1.0e-4      # 8_band.h:61
EPS_REL = 1.0e-4      # 8_band.h:62
LIMIT = 1000          # 8_band.h:63
_EPMACH = np.finfo(np.float64).eps
_UFLOW = np.finfo(np.float64).tiny


def gauss_kronrod(N=30):
    """Nodes (ascending) and weights of the (2N+1)-point Kronrod extension of the N-point
    Gauss-Legendre rule on [-1, 1], and the N-point Gauss weights on the shared nodes (zero on the
    Kronrod-only nodes).

    From the definition: the N+1 added nodes are the zeros of the Stieltjes polynomial E_{N+1},
    the degree-(N+1) polynomial orthogonal to every polynomial of degree <= N with respect to the
    sign-changing weight P_N(x) on [-1, 1].  E_{N+1} is found in the Legendre basis from those N+1
    orthogonality conditions (integrals of degree <= 3N+1, exact with a 2N+2-point Gauss rule), its
    zeros by the Legendre companion matrix polished by Newton steps, and the weights from exactness
    on P_0 .. P_2N (a well-conditioned system in the Legendre basis)."""
    leg = np.polynomial.legendre
    xq, wq = leg.leggauss(2 * N + 2)
    P = np.array([leg.legval(xq, np.eye(N + 2)[j]) for j in range(N + 2)])   # P_j at the nodes
    PN = P[N]
    # M[k][j] = int P_N P_j P_k, k = 0..N, j = 0..N+1; E = sum_j c_j P_j with c_{N+1} = 1
    M = np.einsum("q,kq,jq->kj", wq * PN, P[:N + 1], P)
    c = np.zeros(N + 2)
    c[N + 1] = 1.0
    c[:N + 1] = np.linalg.solve(M[:, :N + 1], -M[:, N + 1])
    z = np.sort(leg.legroots(c).real)
    dc = leg.legder(c)
    for _ in range(3):
        z = z - leg.legval(z, c) / leg.legval(z, dc)
    xg, wg_ = leg.leggauss(N)
    x = np.sort(np.concatenate([xg, z]))
    x = 0.5 * (x - x[::-1])                       # the rule is symmetric
    V = np.array([leg.legval(x, np.eye(2 * N + 1)[k]) for k in range(2 * N + 1)])
    m = np.zeros(2 * N + 1)
    m[0] = 2.0
    wk = np.linalg.solve(V, m)
    wk = 0.5 * (wk + wk[::-1])
    wg = np.zeros_like(x)
    for xi, wi in zip(xg, wg_):
        wg[int(np.argmin(np.abs(x - xi)))] = wi
    return x, wk, wg


_RULE = None


def _rule():
    global _RULE
    if _RULE is None:
        _RULE = gauss_kronrod(30)
    return _RULE


def qk61(f, a, b):
    """DQK61 on [a, b] for a real f taking numpy arrays: (result, abserr, resabs, resasc)."""
    x, wk, wg = _rule()
    c, h = 0.5 * (a + b), 0.5 * (b - a)
    fv = np.asarray(f(c + h * x), dtype=np.float64)
    resk = float(np.dot(wk, fv))
    resg = float(np.dot(wg, fv))
    reskh = 0.5 * resk
    resabs = float(np.dot(wk, np.abs(fv))) * abs(h)
    resasc = float(np.dot(wk, np.abs(fv - reskh))) * abs(h)
    result = resk * h
    abserr = abs((resk - resg) * h)
    if resasc != 0.0 and abserr != 0.0:
        abserr = resasc * min(1.0, (200.0 * abserr / resasc) ** 1.5)
    if resabs > _UFLOW / (50.0 * _EPMACH):
        abserr = max(_EPMACH * 50.0 * resabs, abserr)
    return result, abserr, resabs, resasc


def qag(f, a, b, epsabs=EPS_ABS, epsrel=EPS_REL, limit=LIMIT):
    """DQAGE with the 61-point rule: (result, abserr, ier, intervals used)."""
    result, abserr, resabs, resasc = qk61(f, a, b)
    errbnd = max(epsabs, epsrel * abs(result))
    if abserr <= 50.0 * _EPMACH * resabs and abserr > errbnd:
        return result, abserr, 2, 1
    if (abserr <= errbnd and abserr != resasc) or abserr == 0.0 or limit == 1:
        return result, abserr, (1 if limit == 1 and abserr > errbnd else 0), 1
    lo, hi, area_l, err_l = [a], [b], [result], [abserr]
    area, errsum = result, abserr
    iroff1 = iroff2 = 0
    ier = 0
    for last in range(2, limit + 1):
        i = int(np.argmax(err_l))
        errmax = err_l[i]
        a1, b2 = lo[i], hi[i]
        b1 = 0.5 * (a1 + b2)
        a2 = b1
        area1, error1, _, defab1 = qk61(f, a1, b1)
        area2, error2, _, defab2 = qk61(f, a2, b2)
        area12, erro12 = area1 + area2, error1 + error2
        errsum += erro12 - errmax
        area += area12 - area_l[i]
        if defab1 != error1 and defab2 != error2:
            if abs(area_l[i] - area12) <= 1e-5 * abs(area12) and erro12 >= 0.99 * errmax:
                iroff1 += 1
            if last > 10 and erro12 > errmax:
                iroff2 += 1
        lo[i], hi[i], area_l[i], err_l[i] = a1, b1, area1, error1
        lo.append(a2)
        hi.append(b2)
        area_l.append(area2)
        err_l.append(error2)
        errbnd = max(epsabs, epsrel * abs(area))
        if errsum <= errbnd:
            break
        if iroff1 >= 6 or iroff2 >= 20:
            ier = 2
            break
        if last == limit:
            ier = 1
            break
        if max(abs(a1), abs(b2)) <= (1.0 + 100.0 * _EPMACH) * (abs(a2) + 1000.0 * _UFLOW):
            ier = 3
            break
    return math.fsum(area_l), errsum, ier, len(area_l)


def qag_complex(f, a, b):
    """The reference's complex integrals: real and imaginary parts by two QAG calls."""
    re = qag(lambda x: np.real(f(x)), a, b)[0]
    im = qag(lambda x: np.imag(f(x)), a, b)[0]
    return complex(re, im)
