"""Derive the constants of nbg_libm.h from the mathematics (mpmath, 200 digits;
1600 bits for the 2/pi chunks).

Run:  python gen_libm_consts.py  > consts.txt   (the header embeds the printed values)

* Cody & Waite (1980) additive range reduction: a constant C (pi/2 or ln 2) is
  split into C1 + C2 + C3 where C1 and C2 carry at most `bits` significant bits,
  so that k*C1 and k*C2 are exact for the |k| the reduction admits.
* Polynomial coefficients are the Taylor coefficients 1/k! rounded to double.
* Payne & Hanek (1983) reduction: 2/pi as 24 doubles c_i = M_i * 2^E_i, each
  the 53-bit rounding of 2/pi minus the ones before (mantissa and exponent
  printed apart, since the tail lies far below the double range), and pi/2
  as a double-double.
"""
import mpmath as mp

mp.mp.dps = 200


def to_double_bits(x, bits):
    """Round x to a double with at most `bits` significant bits."""
    if x == 0:
        return mp.mpf(0)
    e = mp.floor(mp.log(abs(x), 2))
    scale = mp.mpf(2) ** (bits - 1 - e)
    return mp.nint(x * scale) / scale


def hexd(x):
    return float(x).hex()


def split(c, bits):
    c1 = to_double_bits(c, bits)
    c2 = to_double_bits(c - c1, bits)
    c3 = to_double_bits(c - c1 - c2, 53)
    return c1, c2, c3


if __name__ == "__main__":
    pio2 = mp.pi / 2
    p1, p2, p3 = split(pio2, 33)
    print("PIO2_1", hexd(p1))
    print("PIO2_2", hexd(p2))
    print("PIO2_3", hexd(p3))
    print("TWO_OVER_PI", hexd(to_double_bits(2 / mp.pi, 53)))
    ln2 = mp.log(2)
    l1, l2, _ = split(ln2, 32)
    l2 = to_double_bits(ln2 - l1, 53)
    print("LN2_1", hexd(l1))
    print("LN2_2", hexd(l2))
    print("LOG2E", hexd(to_double_bits(1 / ln2, 53)))
    for k in range(2, 24):
        print("INV_FACT_%d" % k, hexd(to_double_bits(1 / mp.factorial(k), 53)))
    with mp.workprec(1600):  # the 24 chunks reach 2^-1330
        r = 2 / mp.pi
        for i in range(24):
            e = int(mp.floor(mp.log(abs(r), 2)))
            c = to_double_bits(r, 53)
            print("TWO_OVER_PI_CHUNK %d {%s, %d}" % (i, hexd(c / mp.mpf(2) ** e), e))
            r -= c
    ph = to_double_bits(pio2, 53)
    print("PIO2_HI", hexd(ph))
    print("PIO2_LO", hexd(to_double_bits(pio2 - ph, 53)))
