"""PSRDADA files as the reference's reader expects them (psrdada/psrdada.cpp):
a 4096-byte ASCII header of "KEY value" lines, NUL padded, then the two
polarisations' complex uint16 samples in alternating 2048-sample frames
(pol1 frame, pol2 frame, ...).  Used to write synthetic recordings for the
dual-polarisation ingest and to read recordings back on the host."""
import numpy as np

HEADER_BYTES = 4096
FRAME_SAMPLES = 2048


def write_dada(path, pol1, pol2, bw_mhz, freq_mhz, extra=None):
    """pol1, pol2: uint16 arrays of 2L values (I, Q interleaved), L a multiple
    of 2048.  bw_mhz: bandwidth, freq_mhz: band centre."""
    pol1 = np.ascontiguousarray(pol1, np.uint16).view(np.uint32)
    pol2 = np.ascontiguousarray(pol2, np.uint16).view(np.uint32)
    if pol1.size != pol2.size or pol1.size % FRAME_SAMPLES:
        raise ValueError("both polarisations need the same length, a multiple of 2048 samples")
    keys = {"HDR_VERSION": "1.0", "HDR_SIZE": str(HEADER_BYTES), "BW": repr(float(bw_mhz)),
            "FREQ": repr(float(freq_mhz)), "NBIT": "16", "NCHAN": "1", "NDIM": "2", "NPOL": "2"}
    keys.update(extra or {})
    text = "".join(f"{k:<15} {v}\n" for k, v in keys.items()).encode()
    if len(text) > HEADER_BYTES:
        raise ValueError("header longer than 4096 bytes")
    frames = np.stack([pol1.reshape(-1, FRAME_SAMPLES), pol2.reshape(-1, FRAME_SAMPLES)], axis=1)
    with open(path, "wb") as fh:
        fh.write(text + b"\0" * (HEADER_BYTES - len(text)))
        fh.write(frames.tobytes())


def read_dada(path):
    """(header dict, pol1, pol2) with pol arrays as uint16 of 2L values."""
    with open(path, "rb") as fh:
        head = fh.read(HEADER_BYTES)
        data = np.frombuffer(fh.read(), np.uint32)
    text = head.split(b"\0", 1)[0].decode()
    header = {}
    for line in text.splitlines():
        key, _, val = line.partition(" ")
        if val.strip():
            header[key] = val.strip()
    nfr = data.size // (2 * FRAME_SAMPLES)
    frames = data[:nfr * 2 * FRAME_SAMPLES].reshape(nfr, 2, FRAME_SAMPLES)
    return header, frames[:, 0].reshape(-1).view(np.uint16), frames[:, 1].reshape(-1).view(np.uint16)
