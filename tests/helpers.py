"""Error metrics shared by the parity tests."""
import numpy as np


def db_to_power(db, multiplier=10.0):
    return np.power(10.0, np.asarray(db, np.float64) / multiplier)


def fp32_bar(out_db, ref_db, frames, rel=1e-4, floor=1e-6):
    """Plain-FP32 parity metric on dB outputs [B, M, >=F] (must be <= 1).

    |P_out - P_ref| <= rel * P_ref + floor * max_m P_ref[m, f]: 1e-4 relative in mel power, with a
    floor at 1e-6 of the frame's loudest band.  FP32 rounding in a DFT scales with the frame's
    energy, not the band's, so a band 30+ dB under the frame peak carries more than 1e-4 relative
    error in any FP32 implementation -- RPP HOST's own FFT included (it is off from FP64 by more).
    """
    worst = 0.0
    for b, F in enumerate(frames):
        if F <= 0:
            continue
        po = db_to_power(out_db[b, :, :F])
        pr = db_to_power(ref_db[b, :, :F])
        tol = rel * pr + floor * pr.max(axis=0, keepdims=True)
        worst = max(worst, float((np.abs(po - pr) / tol).max()))
    return worst


def max_rel_power_err(out_db, ref_db, frames):
    return max(float((np.abs(db_to_power(out_db[b, :, :F]) - db_to_power(ref_db[b, :, :F])) /
                      db_to_power(ref_db[b, :, :F])).max()) for b, F in enumerate(frames) if F > 0)


def max_db_err(out, ref, frames):
    return max(float(np.abs(out[b, :, :F].astype(np.float64) - ref[b, :, :F]).max()) for b, F in enumerate(frames)
               if F > 0)
