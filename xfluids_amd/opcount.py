"""Algorithmic FP64 operation counts of the hot-path kernels.

Counts the floating-point operations (add, sub, mul, div, sqrt; fabs and
comparisons excluded) that the plain algorithm needs per characteristic face
flux, by operation class, following the face-flux restatement operation by
operation (Roe average, sound speed, E characteristic fields each with a
left row, the six projected stencil points WENO5 reads, the GLF split, two
WENO5 bodies and the right-column accumulation).  Work the reference does
but never uses (the LLF maximum multiplied by zero, projections of the
stencil points i-3 and i+4) is excluded.
"""


def weno_body_ops(scheme):
    if scheme == 1:  # WENO5-Z body
        s = 12 + 9 + 12          # three smoothness indicators
        w = 1 + 4 * 3 + 2 + 3    # tau5, alpha_k, sum, weights
        r = 3 * (5 + 2) + 2      # three candidate stencils, weighting, sum
        return s + w + r
    # WENO5-JS body (weno5old_BODY)
    s = 3 * (3 + 2 + 3 + 3) + 3  # indicators with the 13/3 forms, + tol
    a = 3 * 4 + 2 + 1 + 3        # products, sum, reciprocal, normalise
    r = 3 * (5 + 1) + 2          # candidate stencils and sum
    return s + a + r


def face_flop_count(E, ncop, scheme):
    """FP64 operations per face for E equations and ncop = Ns - 1."""
    roe = 2 + 2 + 5 * 3 + 2
    if ncop > 0:
        ns = ncop + 1
        c2 = (3 * ns + 3 + 10 + 8 + 2 * 8 + 3 + ncop * (2 * 9 + 3 + 3) + 18 + 5 + 7 + 3 +
              ncop * 6 + 5 + 10 + ncop * 2 + 5 + 5 + ncop * 4 + 3 + 1 + ncop * 4 + 1 + 6 +
              2 * ns * 2)        # + Y from staged U at both cells
    else:
        c2 = 2 + 2
    pre = 5 + 1 + 4 + 1 + 1
    weno = 2 * weno_body_ops(scheme) + 2
    split = 5 * 3 + 5 * 3
    dense = 3 * (3 * E + 6 * E * 2 * 2 + split + weno + 2 * E)
    sparse = (E - 3) * (1 + 6 * 2 * 3 + split + weno + 4)
    return roe + c2 + pre + dense + sparse
