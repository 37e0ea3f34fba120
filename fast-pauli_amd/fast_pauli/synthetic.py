"""Seeded synthetic inputs, portable across hosts.

Every draw comes from a counter-based splitmix64 stream: the i-th 64-bit word
of stream (seed, tag) is splitmix64's i-th output from the state
``mix(seed ^ tag)``.  Integers (letters, supports) are therefore identical on
every machine; normals use Box-Muller on top of them (numpy's log/cos/sin, so
identical wherever numpy is).  The GPU fill kernel (fp_hip_fill_states)
implements the same counter scheme with the device's math library.
"""

from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

TAG_LETTERS = 0x4C455454  # "LETT"
TAG_COEFFS = 0x434F4546   # "COEF"
TAG_STATES = 0x53544154   # "STAT"
TAG_WEIGHTS = 0x57474854  # "WGHT"
TAG_MISC = 0x4D495343     # "MISC"


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def stream_seed(seed: int, tag: int) -> int:
    return int(_mix(np.array([np.uint64(seed) ^ np.uint64(tag)], dtype=np.uint64))[0])


def words(seed: int, start: int, count: int) -> np.ndarray:
    """splitmix64 outputs start..start+count-1 of the generator seeded with `seed`."""
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * GAMMA
    return _mix(z)


def uniform01(seed: int, start: int, count: int) -> np.ndarray:
    w = words(seed, start, count)
    return ((w >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def normals(seed: int, count: int) -> np.ndarray:
    """count N(0,1) draws: pair p uses uniforms (2p, 2p+1); even -> cos, odd -> sin."""
    n_pairs = (count + 1) // 2
    u = uniform01(seed, 0, 2 * n_pairs)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    ang = 2.0 * np.pi * u[1::2]
    out = np.empty(2 * n_pairs, dtype=np.float64)
    out[0::2] = r * np.cos(ang)
    out[1::2] = r * np.sin(ang)
    return out[:count]


def random_codes(n_strings: int, n_qubits: int, seed: int = 18) -> np.ndarray:
    """(n_strings, n_qubits) letter codes, each site iid uniform over {I,X,Y,Z}."""
    w = words(stream_seed(seed, TAG_LETTERS), 0, n_strings * n_qubits)
    return (w >> np.uint64(62)).astype(np.uint8).reshape(n_strings, n_qubits)


def random_codes_max_weight(n_strings: int, n_qubits: int, max_weight: int = 3,
                            seed: int = 18) -> np.ndarray:
    """Strings of weight 1..max_weight: weight uniform, support sites uniform without
    replacement, non-identity letters uniform over {X,Y,Z}."""
    s = stream_seed(seed, TAG_LETTERS)
    per = 2 + 2 * max_weight
    w = words(s, 0, n_strings * per).reshape(n_strings, per)
    out = np.zeros((n_strings, n_qubits), dtype=np.uint8)
    mw = min(max_weight, n_qubits)
    for r in range(n_strings):
        weight = 1 + int(w[r, 0] % np.uint64(mw))
        sites = list(range(n_qubits))
        for k in range(weight):
            idx = int(w[r, 1 + k] % np.uint64(len(sites)))
            site = sites.pop(idx)
            out[r, site] = 1 + int(w[r, 1 + max_weight + k] % np.uint64(3))
    return out


def codes_to_strings(codes: np.ndarray) -> list[str]:
    letters = np.array(list("IXYZ"))
    return ["".join(letters[row]) for row in np.asarray(codes)]


def random_coeffs(n: int, seed: int = 18, dtype=np.complex128) -> np.ndarray:
    z = normals(stream_seed(seed, TAG_COEFFS), 2 * n)
    return (z[0::2] + 1j * z[1::2]).astype(dtype)


def random_states(dim: int, n_states: int, seed: int = 18, normalise: bool = True,
                  dtype=np.complex128) -> np.ndarray:
    """(dim, n_states) row-major states, re/im ~ N(0,1), columns unit 2-norm."""
    z = normals(stream_seed(seed, TAG_STATES), 2 * dim * n_states)
    psi = (z[0::2] + 1j * z[1::2]).reshape(dim, n_states)
    if normalise:
        psi /= np.sqrt((np.abs(psi) ** 2).sum(axis=0, keepdims=True))
    return np.ascontiguousarray(psi.astype(dtype))


def random_real_matrix(rows: int, cols: int, seed: int = 18) -> np.ndarray:
    return normals(stream_seed(seed, TAG_WEIGHTS), rows * cols).reshape(rows, cols)


def gpu_seed(seed: int) -> int:
    """Seed passed to fp_hip_fill_states for the same logical state stream."""
    return stream_seed(seed, TAG_STATES)
