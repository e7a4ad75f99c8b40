"""Independent pure-Python (big-integer) arithmetic used to pin the oracle.

Nothing here calls the oracle or the engine: these are textbook definitions
(direct polynomial evaluation, CRT composition, schoolbook products) evaluated
with Python integers, so a transposed index, a dropped term or a wrong sign in
the oracle makes the comparisons fail.
"""
from __future__ import annotations


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
    for p in small:
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def primes_descending(N: int, bits: int, count: int) -> list[int]:
    """Primes p = 1 mod 2N below 2^bits, largest first."""
    out, v = [], (1 << bits) - 2 * N + 1
    while len(out) < count:
        assert v > (1 << (bits - 1))
        if is_prime(v):
            out.append(v)
        v -= 2 * N
    return out


def coeff_modulus(N: int, bit_sizes: list[int]) -> list[int]:
    """CoeffModulus::Create: each size class is consumed from its smallest prime."""
    pools = {b: primes_descending(N, b, bit_sizes.count(b)) for b in set(bit_sizes)}
    return [pools[b].pop() for b in bit_sizes]


def brv(x: int, bits: int) -> int:
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def ntt_direct(a: list[int], psi: int, q: int) -> list[int]:
    """Slot j = a(psi^(2 brv(j) + 1)) (negacyclic evaluation, bit-reversed order)."""
    N = len(a)
    logN = N.bit_length() - 1
    out = []
    for j in range(N):
        x = pow(psi, 2 * brv(j, logN) + 1, q)
        acc, p = 0, 1
        for c in a:
            acc += c * p
            p = p * x % q
        out.append(acc % q)
    return out


def negacyclic_mul(a: list[int], b: list[int], q: int) -> list[int]:
    N = len(a)
    out = [0] * N
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            if k < N:
                out[k] += x * y
            else:
                out[k - N] -= x * y
    return [v % q for v in out]


def automorphism(a: list[int], elt: int, q: int) -> list[int]:
    """a(X) -> a(X^elt) mod X^N + 1."""
    N = len(a)
    out = [0] * N
    for i, c in enumerate(a):
        k = i * elt % (2 * N)
        if k < N:
            out[k] = (out[k] + c) % q
        else:
            out[k - N] = (out[k - N] - c) % q
    return out


def crt(residues: list[int], moduli: list[int]) -> int:
    Q = 1
    for m in moduli:
        Q *= m
    x = 0
    for r, m in zip(residues, moduli):
        h = Q // m
        x += r * h * pow(h, -1, m)
    return x % Q
