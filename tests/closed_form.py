"""Independent closed-form brute force of D(x, y, z, k) and S2_hard(x, y, z, c).

This does not follow the segmented-sieve structure of the reference or of
the oracle: it enumerates the leaf set from its definition and evaluates
phi(n, a) by Legendre's recursion.  Tiny inputs only.

  D(x,y,z,k) = sum_{b=k+1}^{pi(sqrt z)} sum_{m} -mu(m) phi(x/(p_b m), b-1)
             + sum_{b=pi(sqrt z)+1}^{pi(x*)} sum_{q} phi(x/(p_b q), b-1)
  with m coprime to 2*3*5*7*11, z/p_b < m <= min(x/p_b^3, z), mu(m) != 0,
  lpf(m) > p_b, mpf(m) <= y;  q prime, p_b < q <= min(x/p_b^3, y);
  every leaf value n = x/(p_b m) lies below the sieve limit x/z.
  (Limits as src/gourdon/D.cpp:122-133, 279-290.)

  S2_hard(x,y,z,c) = sum_{b=c+1}^{pi(sqrt y)} sum_{m} -mu(m) phi(x/(p_b m), b-1)
                   + sum_{b=pi(sqrt y)+1}^{pi(min(sqrt z, y))} sum_{q} phi(x/(p_b q), b-1)
  with m coprime to 2310, y/p_b < m <= y, mu(m) != 0, lpf(m) > p_b;
  q prime, p_b < q <= min(y, z/p_b); leaf values below z.
  (src/deleglise-rivat/S2_hard.cpp:98-155.)
"""
from functools import lru_cache
import math


def primes_upto(n):
    n = max(n, 2)
    s = bytearray([1]) * (n + 1)
    s[0] = s[1] = 0
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = bytearray(len(s[i * i::i]))
    return [i for i in range(n + 1) if s[i]]


def factor_info(n_max):
    lpf = [0] * (n_max + 1)
    mpf = [0] * (n_max + 1)
    mu = [1] * (n_max + 1)
    for p in range(2, n_max + 1):
        if lpf[p] == 0:
            for m in range(p, n_max + 1, p):
                if lpf[m] == 0:
                    lpf[m] = p
                mpf[m] = p
                mu[m] = -mu[m]
            for m in range(p * p, n_max + 1, p * p):
                mu[m] = 0
    return lpf, mpf, mu


def make_phi(primes):
    @lru_cache(maxsize=None)
    def phi(n, a):
        if a == 0 or n == 0:
            return n
        if n < primes[a - 1]:
            return 1 if n >= 1 else 0
        return phi(n, a - 1) - phi(n // primes[a - 1], a - 1)
    return phi


def isqrt(n):
    return math.isqrt(n)


def x_star(x, y):
    y = max(y, 1)
    xs = max(round_root4(x), -(-x // (y * y)))
    xs = min(xs, y, isqrt(x // y))
    return max(xs, 1)


def round_root4(x):
    return isqrt(isqrt(x))


def coprime2310(m):
    return m % 2 and m % 3 and m % 5 and m % 7 and m % 11


def D(x, y, z, k):
    P = primes_upto(max(y, isqrt(z)) + 2)
    lpf, mpf, mu = factor_info(max(z, 2))
    phi = make_phi(tuple(P))
    pi = lambda n: sum(1 for p in P if p <= n)
    xz = x // z
    xs = x_star(x, y)
    total = 0
    b_split = pi(isqrt(z))
    for b in range(k + 1, pi(xs) + 1):
        p = P[b - 1]
        if b <= b_split:
            for m in range(z // p + 1, min(x // p ** 3, z) + 1):
                if not coprime2310(m) or mu[m] == 0 or lpf[m] <= p or mpf[m] > y:
                    continue
                n = x // (p * m)
                if n < xz:
                    total -= mu[m] * phi(n, b - 1)
        else:
            for q in P:
                if q <= p:
                    continue
                if q > min(x // p ** 3, y):
                    break
                n = x // (p * q)
                if n < xz:
                    total += phi(n, b - 1)
    return total


def S2_hard(x, y, z, c):
    P = primes_upto(max(y, isqrt(z)) + 2)
    lpf, mpf, mu = factor_info(max(y, 2))
    phi = make_phi(tuple(P))
    pi = lambda n: sum(1 for p in P if p <= n)
    total = 0
    b_split = pi(isqrt(y))
    b_top = pi(min(isqrt(z), y)) if z > y else b_split
    for b in range(c + 1, b_top + 1):
        p = P[b - 1]
        if b <= b_split:
            for m in range(y // p + 1, y + 1):
                if not coprime2310(m) or mu[m] == 0 or lpf[m] <= p:
                    continue
                n = x // (p * m)
                if n < z:
                    total -= mu[m] * phi(n, b - 1)
        else:
            for q in P:
                if q <= p:
                    continue
                if q > min(y, z // p):
                    break
                n = x // (p * q)
                if n < z:
                    total += phi(n, b - 1)
    return total
