"""Write tests/golden/fixtures.json from first principles (pure Python).

No oracle or engine code is involved: primes by enumeration, NTT outputs by
direct evaluation at psi^(2 brv(j) + 1), Galois slot permutations from the
evaluation points (slot j of sigma_elt(a) reads slot k with
2 brv(k) + 1 = (2 brv(j) + 1) * elt mod 2N), rescale by CRT floor division.
Run: python tests/golden/make_fixtures.py
"""
import json
import os
import random
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from tests import _pyref as P  # noqa: E402


def min_root(q, N):
    x = 2
    while pow(x, N, q) != q - 1:
        x += 1
    return x


def main():
    rnd = random.Random(2024)
    fx = {"primes": [], "ntt": [], "galois": [], "rescale": []}
    for logN, bits, nP in [(13, [50] * 8 + [60], 1), (12, [60, 40, 40, 60], 1), (16, [50] * 24 + [60], 1),
                           (16, [51] + [46] * 20 + [51] * 14 + [51] * 5, 5)]:
        fx["primes"].append({"logN": logN, "bits": bits, "nP": nP, "primes": P.coeff_modulus(1 << logN, bits)})
    logN, bits = 4, [17, 18, 19]
    N = 1 << logN
    primes = P.coeff_modulus(N, bits)
    for k, q in enumerate(primes):
        psi = min_root(q, N)
        a = [rnd.randrange(q) for _ in range(N)]
        fx["ntt"].append({"logN": logN, "bits": bits, "prime_index": k, "input": a,
                          "output": P.ntt_direct(a, psi, q)})
    for step in (1, -1, 2, 0):
        m = 2 * N
        elt = m - 1 if step == 0 else pow(5, step % (N // 2), m)
        pts = {(2 * P.brv(k, logN) + 1) % m: k for k in range(N)}
        table = [pts[(2 * P.brv(j, logN) + 1) * elt % m] for j in range(N)]
        fx["galois"].append({"logN": logN, "bits": bits, "elt": elt, "table": table})
    Ql = 3
    coeff = [[rnd.randrange(primes[i]) for _ in range(N)] for i in range(Ql)]
    out = [[0] * N for _ in range(Ql - 1)]
    for j in range(N):
        C = P.crt([coeff[i][j] for i in range(Ql)], primes[:Ql])
        f = C // primes[Ql - 1]
        for i in range(Ql - 1):
            out[i][j] = f % primes[i]
    fx["rescale"].append({"logN": logN, "bits": bits, "Ql": Ql, "coeff_in": sum(coeff, []),
                          "coeff_out": sum(out, [])})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
    with open(path, "w") as f:
        json.dump(fx, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
