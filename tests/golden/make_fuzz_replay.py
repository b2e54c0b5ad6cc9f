"""Writes tests/golden/fuzz_replay.npz from the reference's fuzz corpus.

Run in the container that holds the reference, after `make -C oracle fuzz`
(which compiles the reference's fuzz harness, fuzz/harness.h, with its CPU
sources into oracle/_ref/libtsfuzz.so).  The corpus
(reference fuzz/corpus/{parse,mutation}, replayed there by fuzz/replay.cpp:23-44)
is data:

  parse     each file is an arbitrary byte string given to the parser as is
  mutation  each file is a recipe (image, pattern, seed, then up to eight
            12-byte mutation commands) that the harness applies to a canonical
            archive (fuzz/harness.h:258-339); the fixture keeps the mutated
            archive the harness parsed, not the recipe

For every input the fixture holds the bytes given to the parser, the reference
parser's status under the harness's limits (fuzz/harness.h:25: 1 MiB file,
32768 tiles, 8 MiB metadata) and whether the harness's own checks passed on
it (decode, padding, canonical re-encode no larger, round trip).
"""
import ctypes as C
import os
import sys

import numpy as np

CORPUS = "/root/reference/fuzz/corpus"
HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libtsfuzz.so")
OUT = os.path.join(HERE, "fuzz_replay.npz")


def main():
    lib = C.CDLL(LIB)
    u8p = C.POINTER(C.c_uint8)
    lib.ref_fuzz_check_file.argtypes = [u8p, C.c_uint64, C.POINTER(C.c_int), C.c_char_p, C.c_uint64]
    lib.ref_fuzz_mutation.argtypes = [u8p, C.c_uint64, C.POINTER(C.c_int), u8p, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.c_char_p, C.c_uint64]
    names, kinds, statuses, clean, blobs = [], [], [], [], []
    out = np.zeros(4 << 20, np.uint8)
    for kind, sub in enumerate(("parse", "mutation")):
        folder = os.path.join(CORPUS, sub)
        for name in sorted(os.listdir(folder)):
            data = np.fromfile(os.path.join(folder, name), np.uint8)
            ptr = data.ctypes.data_as(u8p) if data.size else None
            status, problem = C.c_int(0), C.create_string_buffer(256)
            if kind == 0:
                bad = lib.ref_fuzz_check_file(ptr, data.size, C.byref(status), problem, 256)
                file = data.tobytes()
            else:
                written = C.c_uint64(0)
                bad = lib.ref_fuzz_mutation(ptr, data.size, C.byref(status), out.ctypes.data_as(u8p),
                                            out.size, C.byref(written), problem, 256)
                assert written.value <= out.size, name
                file = out[:written.value].tobytes()
            if bad:
                print(f"{sub}/{name}: harness reports {problem.value.decode()}", file=sys.stderr)
            names.append(f"{sub}/{name}")
            kinds.append(kind)
            statuses.append(status.value)
            clean.append(not bad)
            blobs.append(file)
    offsets = np.zeros(len(blobs) + 1, np.int64)
    offsets[1:] = np.cumsum([len(b) for b in blobs])
    np.savez_compressed(OUT, data=np.frombuffer(b"".join(blobs), np.uint8), offsets=offsets,
                        kind=np.array(kinds, np.uint8), status=np.array(statuses, np.int8),
                        harness_clean=np.array(clean, bool), name=np.array(names))
    accepted = sum(1 for s in statuses if s == 0)
    print(f"wrote {len(blobs)} inputs ({offsets[-1]} bytes, {accepted} accepted) to {OUT}")


if __name__ == "__main__":
    main()
