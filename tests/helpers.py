"""Shared test helpers: run GPU operations through the C-ABI on numpy inputs."""
from __future__ import annotations

import numpy as np

import tilestream_amd as ts
from tilestream_amd import api
from tilestream_amd._native import BatchItem, CompressedInput, ConstBytes, MutableBytes, Roi, TileMapping

SENTINEL = 0xA5


class DeviceOutput:
    """A pitched device output with sentinel-filled padding."""

    def __init__(self, ptype, width, height, pad_elems=3):
        self.ptype = ptype
        self.width, self.height = width, height
        self.s = api.sample_bytes(ptype)
        self.pitch = (width + pad_elems) * self.s
        self.mem = api.DeviceMemory(self.pitch * height + 64)
        self.mem.fill(SENTINEL)

    def view(self):
        return api.device_view(self.mem, self.ptype, self.width, self.height, self.pitch)

    def read(self):
        raw = self.mem.download()
        rows = raw[: self.pitch * self.height].reshape(self.height, self.pitch)
        pad = rows[:, self.width * self.s:]
        assert (pad == SENTINEL).all(), "row padding was written"
        assert (raw[self.pitch * self.height:] == SENTINEL).all(), "bytes past the image were written"
        body = np.ascontiguousarray(rows[:, : self.width * self.s])
        return body.view(np.uint16 if self.s == 2 else np.uint32).reshape(self.height, self.width)


def upload_archive(data: bytes, offset: int = 0, extra: int = 0):
    mem = api.DeviceMemory(offset + len(data) + extra)
    mem.fill(0xFF)
    mem.upload(data, offset)
    return mem


class Job:
    """One GPU batch: plan + buffers, kept alive together."""

    def __init__(self):
        self.keep = []
        self.items = []
        self.outputs = []

    def add_decode(self, data: bytes, roi=None, calibrated=False, offset=None, gain=None,
                   mem=None, mem_offset=0, metadata=None, mappings=None):
        st, meta = api.parse_metadata(data) if metadata is None else (ts.Status.success, metadata)
        assert st == ts.Status.success, st
        summ = meta.summary
        img = summ.image
        if mem is None:
            mem = upload_archive(data)
        x, y, w, h = roi if roi is not None else (0, 0, img.width, img.height)
        out = DeviceOutput(ts.PixelType.float32 if calibrated else img.type, w, h)
        item = BatchItem()
        item.output = out.view()
        item.compressed_input = CompressedInput()
        item.compressed_input.bytes = ConstBytes(mem.ptr + mem_offset, mem.size - mem_offset)
        item.compressed_input.metadata = meta.handle
        if mappings is not None:
            arr = (TileMapping * len(mappings))(*[TileMapping(a, b) for a, b in mappings])
            item.compressed_input.mappings = arr
            item.compressed_input.mapping_count = len(mappings)
            self.keep.append(arr)
        item.roi = Roi(x, y, w, h)
        if calibrated:
            item.calibration.offset = self._param(offset, w, h, 0.0)
            item.calibration.gain = self._param(gain, w, h, 1.0)
        self.keep += [meta, mem]
        self.items.append(item)
        self.outputs.append(out)
        return out

    def _param(self, value, w, h, default):
        if value is None:
            return api.scalar(default)
        if isinstance(value, np.ndarray):
            m = api.DeviceMemory.from_host(np.ascontiguousarray(value, np.float32))
            self.keep.append(m)
            return api.map_parameter(api.device_view(m, ts.PixelType.float32, w, h, w * 4, const=True))
        return api.scalar(float(value))

    def run(self, operation, shape=None, workspace_extra=0):
        st, plan = api.prepare_batch(operation, self.items)
        assert st == ts.Status.success, st
        if shape is not None:
            assert plan.set_decode_shape(*shape) == ts.Status.success, shape
        ws = api.DeviceMemory(plan.workspace_bytes() + workspace_extra) if plan.workspace_bytes() else None
        results = api.DeviceMemory(16 * len(self.items))
        results.fill(0x5A)
        fn = {ts.Operation.decompress: api.decompress_async, ts.Operation.decode_roi: api.decode_roi_async,
              ts.Operation.decode_calibrate: api.decode_calibrate_async}[operation]
        st = fn(plan, ws, results, len(self.items), None)
        assert st == ts.Status.success, st
        assert api.synchronize(None) == ts.Status.success
        status, written = api.read_results(results, len(self.items))
        self.keep += [plan, ws, results]
        return status, written, [o.read() for o in self.outputs]


def decode_on_gpu(data: bytes, roi=None, shape=None):
    job = Job()
    job.add_decode(data, roi)
    op = ts.Operation.decode_roi if roi is not None else ts.Operation.decompress
    status, written, outs = job.run(op, shape)
    assert status[0] == 0, status
    return outs[0], int(written[0])
