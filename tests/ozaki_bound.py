"""Error bound of the Ozaki split with signed-byte limbs (the scheme of
skinny_dtsmttsm_ozaki), used by the oracle and GPU tests.

An element x with column scale 2^E (|x| < 2^E) is rounded to
q = rint(x 2^(8S - 2 - E)) and written in S signed bytes d_0 .. d_(S-1)
(|d_0| <= 64, the others <= 128).  A product then carries
  * the rounding of both operands: 2 * 2^-(8S - 1) of 2^(Ea + Eb);
  * the dropped levels g = s + t >= L (L = 3 for 3 limbs, S + 1 for 4 and 5):
    each pair of lower digits at most 2^14, weighted 2^(-8g - 12) of
    2^(Ea + Eb) (x y = 2^(Ea + Eb - 12) sum_(s,t) d_s d'_t 2^-8(s + t)).
The scale is within a factor 2 of the chunk's largest |x|, and the rounding
errors of the K products are independent, so the error of a whole product
against sum_k |a_k| |b_k| stays below ozaki_bound(S) by a wide margin.
"""


def ozaki_levels(S):
    return 3 if S == 3 else S + 1


def ozaki_bound(S):
    L = ozaki_levels(S)

    def npairs(g):
        return sum(1 for s in range(S) if 0 <= g - s < S)

    dropped = sum(npairs(g) * 2.0 ** (14 - 8 * g - 12) for g in range(L, 2 * S - 1))
    return 2 * 2.0 ** -(8 * S - 1) + dropped
