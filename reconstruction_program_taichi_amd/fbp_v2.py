"""View-block streaming FBP: drop-in for run_mgfbp_v2.py (run_mgfbp_v2, Mgfbp_v2).

The reference (run_mgfbp_v2.py:24-58, 61-103) keeps the sinogram in host
memory and moves ONE view at a time to the GPU: BH correction, weighting and
filtering of that view, then a back-projection that adds it into the volume
(run_mgfbp_v2.py:67-92).  Device memory is bounded by the volume plus one
view, which is what lets it reconstruct scans larger than the GPU's memory.

Here the same bound holds for blocks of views (config key ViewsPerBlock;
default: the largest block whose device buffers fit in 1 GiB):
  * the file is read once into pinned host memory;
  * block b+1 is copied to the device on a copy stream (one 2-D DMA,
    mct_copy_views) while block b is weighted (mct_weight_views), filtered,
    packed and back-projected (mct_backproject_views) on the compute stream;
    two device slots alternate, ordered by events;
  * each block's back-projection adds to the volume voxel by voxel in view
    order, so every voxel's sum is formed in the same order as the resident
    path's (Mgfbp.reconstruct), and the last block clears the truncation
    marks (SetTruncatedRegionToZero, run_mgfbp_v2.py:232-236).

The result equals run_mgfbp's bit for bit: the streamed blocks run the same
kernels as the resident path, and a voxel that leaves the detector in some
view ends as 0 in both.  The one difference the reference itself has: a
voxel whose running sum falls below -5000 is cleared to 0 here as there.
"""
from __future__ import annotations

import gc
import os
import re
import sys
import time
import warnings

import numpy as np
import torch

from . import _abi
from .config import ReadConfigFile, fail
from .device import DTYPES, LAYOUTS
from .fbp import Mgfbp

# Device bytes per view block when ViewsPerBlock is not given.
DEFAULT_BLOCK_BYTES = 1 << 30


def run_mgfbp_v2(file_path):
    """run_mgfbp_v2.py:24-58: reconstruct every matching file, streaming the
    sinogram through the GPU in blocks of views; returns the last
    reconstruction."""
    print("Performing FBP from MandoCT (MI355X HIP, view-block streaming) ...")
    start = time.time()
    warnings.filterwarnings("ignore", category=UserWarning)
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    config_dict = ReadConfigFile(file_path)
    fbp = Mgfbp_v2(config_dict)
    img_recon = fbp.MainFunction()
    elapsed = time.time() - start
    if fbp.file_processed_count > 0:
        print(f"\nA total of {fbp.file_processed_count:d} file(s) are reconstructed!")
        print(f"Time cost is {elapsed:.3} sec\n")
    else:
        print(f"\nWarning: Did not find files like {fbp.input_files_pattern:s} in {fbp.input_dir:s}.")
        print("No images are reconstructed!\n")
    del fbp
    gc.collect()
    return img_recon


class Mgfbp_v2(Mgfbp):
    """run_mgfbp_v2.py:61: Mgfbp whose MainFunction streams views."""

    def __init__(self, config_dict, device: int = 0, ctx=None):
        super().__init__(config_dict, device, ctx)
        vpb = config_dict.get("ViewsPerBlock", 0)
        if isinstance(vpb, bool) or not isinstance(vpb, int) or vpb < 0:
            fail("ERROR: ViewsPerBlock must be a non-negative integer!")
        self.views_per_block = int(vpb) or self.default_views_per_block()
        self.host_sgm = None

    def default_views_per_block(self) -> int:
        """Views whose device buffers (two upload slots, the filtered block,
        the vertical filter's output, the packed taps) fit in
        DEFAULT_BLOCK_BYTES."""
        rows, nu = self.det_elem_count_vertical_actual, self.det_elem_count_horizontal
        plain = 4 * rows * nu
        packed = _abi.lib().mct_packed_bytes(rows, 1, nu, DTYPES[self.storage],
                                             LAYOUTS[self.tap_layout()])
        per_view = 3 * plain + (plain if self.apply_gauss_vertical else 0) + max(int(packed), 0)
        return max(1, min(self.view_num, DEFAULT_BLOCK_BYTES // per_view))

    # ------------------------------------------------------------ files
    def load_rows(self) -> None:
        """The file's recon-range rows into pinned host memory, self.host_sgm
        (rows, sgm_height, n_u), so blocks of views can be copied to the
        device while kernels run; a sinogram file is read straight into it."""
        nu, rows = self.det_elem_count_horizontal, self.det_elem_count_vertical_actual
        self.img_sgm = self.host_sgm = None
        host = torch.empty((rows, self.sgm_height, nu), dtype=torch.float32, pin_memory=True)
        buf = host.numpy()
        if self.input_file_form == "sinogram":
            off = self.det_elem_vertical_recon_range_begin * 4 * self.sgm_height * nu
            with open(self.input_path, "rb") as fh:
                fh.seek(off)
                got = fh.readinto(memoryview(buf).cast("B"))
            if got != buf.nbytes:
                raise ValueError(f"{self.input_path}: {got} bytes where {buf.nbytes} were expected")
        else:
            buf[...] = self.read_file_rows()
        self.host_sgm = host

    # ------------------------------------------------------------ device
    def reconstruct_streamed(self, host: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """FBP of the first view_num views of the host FP32 sinogram
        (rows, views, n_u), streamed in blocks of views_per_block views; out is
        zeroed first (as the reference's volume is before each file,
        run_mgfbp.py:1302-1310).  host should be pinned for the uploads to
        overlap kernels; it must stay alive until the returned volume has
        been read (the uploads run on a stream of their own)."""
        ctx, plan = self.ctx(), self.plan()
        dev = torch.device("cuda", self._device)
        rows, nfile, nu = host.shape
        if rows != self.det_elem_count_vertical_actual or nu != self.det_elem_count_horizontal:
            raise ValueError(f"host sinogram shape {tuple(host.shape)} does not match the geometry")
        if nfile < self.view_num:
            raise ValueError(f"host sinogram has {nfile} views, the scan {self.view_num}")
        nviews = self.view_num
        block = max(1, min(self.views_per_block, nviews))
        blocks = [(j0, min(block, nviews - j0)) for j0 in range(0, nviews, block)]
        compute = torch.cuda.current_stream(dev)
        ctx.bind_stream(compute)
        copy = torch.cuda.Stream(dev)
        layout = self.tap_layout()
        pbytes = lambda nb: int(_abi.lib().mct_packed_bytes(rows, nb, nu, DTYPES[self.storage],
                                                            LAYOUTS[layout]))
        slots = [torch.empty(rows * block * nu, dtype=torch.float32, device=dev) for _ in range(2)]
        loaded = [torch.cuda.Event() for _ in range(2)]
        consumed = [None, None]
        packed = torch.empty(pbytes(block), dtype=torch.uint8, device=dev)
        if out is None:
            out = torch.empty((self.img_dim_z, self.img_dim, self.img_dim), dtype=torch.float32,
                              device=dev)
        out.zero_()
        # the slots may reuse memory that earlier compute-stream work still reads
        copy.wait_stream(compute)

        def upload(b):
            j0, nb = blocks[b]
            s = b % 2
            blk = slots[s][:rows * nb * nu].view(rows, nb, nu)
            if consumed[s] is not None:  # the block two back has been read
                copy.wait_event(consumed[s])
            ctx.copy_views(host, j0, nb, blk, stream=copy)
            loaded[s].record(copy)
            return blk

        fused = self.fused_filter_pack()
        taps = self.recon_kernel() if fused else None
        bh = self.array_bh_coefficients if self.bool_bh_correction else None
        nxt = upload(0)
        filt = None
        for b, (j0, nb) in enumerate(blocks):
            blk = nxt
            if b + 1 < len(blocks):
                nxt = upload(b + 1)
            compute.wait_event(loaded[b % 2])
            pk = packed[:pbytes(nb)]
            if fused:  # the MFMA filter writes the block's ROWS taps itself
                plan.weight_filter_pack(blk, pk, taps, self.det_elem_width, self.filter_method, j0, bh)
            else:
                filt = self.weight_and_filter(blk, j0)
                ctx.pack_into(filt, self.storage, layout, pk)
            plan.backproject_views(pk, self.storage, layout, j0, nb, out,
                                   finish=b == len(blocks) - 1)
            ev = torch.cuda.Event()
            ev.record(compute)
            consumed[b % 2] = ev
        # Mgfbp_v2 keeps one view on the device: its filtered-sinogram file
        # holds the last view only (run_mgfbp_v2.py:93, 108-109)
        self.last_filtered = filt[:, -1:, :] if filt is not None else None
        return out

    def MainFunction(self):
        """run_mgfbp_v2.py:62-94."""
        self.file_processed_count = 0
        for file in os.listdir(self.input_dir):
            if re.match(self.input_files_pattern, file) and self.ReadSinogram(file):
                self.file_processed_count += 1
                print("Reconstructing %s ..." % self.input_path)
                vol = self.reconstruct_streamed(self.host_sgm)
                self.img_recon = vol.cpu().numpy()
                print("\nSaving to %s !" % self.output_path)
                self.SaveReconImg()
                self.SaveFilteredSinogram()
        return self.img_recon
