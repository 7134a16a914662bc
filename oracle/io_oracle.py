"""Test infrastructure: numpy restatement of SenseMap's normal-map encoding,
the checker for pm_normal_encode / pm_normal_decode and the device encoder.

NormalMap::Encode (src/mvs/normal_map.cc:155-176) and Decode (:133-152),
with decode_factor = 32767.f (:8).  numpy float32 arithmetic rounds every
operation to float, as the reference's SSE code does; the double steps of
the reference are restated in float64 where they are not exact in float."""
import numpy as np

F = np.float32(32767.0)
INV = np.float32(1.0) / F


def encode(normal):
    n0, n1, n2 = (np.asarray(normal[k], np.float32) for k in range(3))
    with np.errstate(divide="ignore", invalid="ignore"):
        # float s = (1 - normal[0]) * 0.5;  -- (1 - n0) in float, * 0.5 in double, back to float
        s = ((np.float32(1) - n0).astype(np.float64) * 0.5).astype(np.float32)
        # float m = 1.0f / std::sqrt(1 - s);
        m = np.float32(1) / np.sqrt(np.float32(1) - s)
        # float q1 = normal[1] * 0.5 * m;  -- double products, rounded to float once
        q1 = (n1.astype(np.float64) * 0.5 * m.astype(np.float64)).astype(np.float32)
        q2 = (n2.astype(np.float64) * 0.5 * m.astype(np.float64)).astype(np.float32)
    lo, hi = np.float32(-0.999999), np.float32(0.999999)
    # std::max(q, lo) keeps NaN; std::min(hi, NaN) gives hi
    q1 = np.where(q1 < lo, lo, q1)
    q2 = np.where(q2 < lo, lo, q2)
    q1 = np.where(q1 < hi, q1, hi)
    q2 = np.where(q2 < hi, q2, hi)
    iq1 = ((q1 + np.float32(1)) * F).astype(np.uint32)
    iq2 = ((q2 + np.float32(1)) * F).astype(np.uint32)
    return (iq1 << np.uint32(16)) | iq2


def decode(codes):
    c = np.asarray(codes, np.uint32)
    q1 = (c >> np.uint32(16)).astype(np.float32) * INV - np.float32(1)
    q2 = (c & np.uint32(0xFFFF)).astype(np.float32) * INV - np.float32(1)
    s = q1 * q1 + q2 * q2
    r = np.sqrt(np.float32(1) - s)
    n0 = (1.0 - 2.0 * s.astype(np.float64)).astype(np.float32)
    return np.stack([n0, (np.float32(2) * q1) * r, (np.float32(2) * q2) * r])
