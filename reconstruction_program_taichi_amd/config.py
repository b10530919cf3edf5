"""JSONC config files and the reference's config helpers.

remove_comments / load_jsonc / save_jsonc / ReadConfigFile / imreadRaw /
XuLeastSquareSolution behave as the reference's module-level helpers
(run_mgfbp.py:1313-1359): comments are stripped with the regex
``//.*?$|/\\*.*?\\*/`` (multiline, dotall) - which also strips ``//`` inside
strings, as the reference does - and the rest is parsed with json.loads.
Config errors print the reference's message and raise SystemExit.
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

PI = 3.1415926536  # run_mgfbp.py:16


def remove_comments(jsonc_str: str) -> str:
    pattern = re.compile(r"//.*?$|/\*.*?\*/", re.MULTILINE | re.DOTALL)
    return re.sub(pattern, "", jsonc_str)


def save_jsonc(save_path: str, data) -> None:
    assert save_path.split(".")[-1] == "jsonc"
    with open(save_path, "w") as f:
        json.dump(data, f)


def load_jsonc(file_path: str):
    with open(file_path, "r") as f:
        return json.loads(remove_comments(f.read()))


def ReadConfigFile(file_path: str):
    return load_jsonc(file_path)


def imreadRaw(path: str, height: int, width: int, dtype=np.float32, nSlice: int = 1,
              offset: int = 0, gap: int = 0) -> np.ndarray:
    """run_mgfbp.py:1331-1342."""
    with open(path, "rb") as fp:
        fp.seek(offset)
        if gap == 0:
            return np.frombuffer(fp.read(), dtype=dtype, count=nSlice * height * width).reshape(
                (nSlice, height, width)).copy()
        nbytes = height * width * np.dtype(dtype).itemsize
        arr = np.zeros((nSlice, height, width), dtype=dtype)
        for i in range(nSlice):
            arr[i] = np.frombuffer(fp.read(nbytes), dtype=dtype).reshape((height, width))
            fp.seek(gap, os.SEEK_CUR)
        return arr


def imaddRaw(img, path: str, dtype=None, idx=1) -> None:
    """run_mgfpj.py:708-722: write the raw bytes of img to path when idx is 0,
    append them otherwise (converted to dtype first when given)."""
    if dtype is not None:
        img = img.astype(dtype)
    with open(path, "wb" if idx == 0 else "ab") as fp:
        fp.write(np.ascontiguousarray(img).tobytes())


def imwriteRaw(img: np.ndarray, path: str, dtype=np.float32) -> None:
    """Raw C-order bytes (the reference uses crip.io.imwriteRaw)."""
    np.ascontiguousarray(img, dtype=dtype).tofile(path)


def imwriteTiff(img: np.ndarray, path: str, dtype=np.float32) -> None:
    """Multi-page float TIFF (the reference uses crip.io.imwriteTiff)."""
    from PIL import Image

    img = np.asarray(img, dtype=dtype)
    if img.ndim == 2:
        img = img[None]
    pages = [Image.fromarray(np.ascontiguousarray(p)) for p in img]
    pages[0].save(path, save_all=True, append_images=pages[1:])


def XuLeastSquareSolution(matrix_B, vec_y):
    """x = (B^T B)^-1 B^T y (run_mgfbp.py:1355-1359)."""
    temp = np.matmul(np.linalg.inv(np.matmul(matrix_B.transpose(), matrix_B)), matrix_B.transpose())
    return np.matmul(temp, vec_y)


def fail(msg: str):
    """The reference's config-error convention: print, then sys.exit()."""
    print(msg)
    sys.exit()


def is_num(x) -> bool:
    return isinstance(x, (int, float)) and not isinstance(x, bool)


def first_key(cfg: dict, *names):
    for n in names:
        if n in cfg:
            return n, cfg[n]
    return None, None
