"""Image I/O for the command line (fpie/io.py:10-39), on Pillow.

Images are read as H x W x 3 uint8 arrays in BGR channel order, the order
cv2.imread gives the reference, so per-channel outputs (err) print in the
same order; gray images are stacked to 3 channels and alpha is dropped.
"""

from __future__ import annotations

import os
import warnings

import numpy as np


def read_image(name: str) -> np.ndarray:
  from PIL import Image

  if not os.path.exists(name):
    raise FileNotFoundError(f"Failed to read image: {name}")
  img = np.asarray(Image.open(name))
  if img.ndim == 2:
    img = np.stack([img, img, img], axis=-1)
  elif img.ndim == 3 and img.shape[-1] == 4:
    img = img[..., :3]
  return np.ascontiguousarray(img[..., ::-1])  # RGB -> BGR


def write_image(name: str, image: np.ndarray) -> None:
  from PIL import Image

  Image.fromarray(np.ascontiguousarray(image[..., ::-1])).save(name)


def read_images(src_name: str, mask_name: str, tgt_name: str):
  """(source, mask, target) of the command line.  A missing mask file means
  "blend the whole source": a warning, then an all-255 mask of the source's
  shape (the behaviour of fpie/io.py:27-39)."""
  images = {"src": read_image(src_name), "tgt": read_image(tgt_name)}
  if not os.path.exists(mask_name):
    warnings.warn("No mask file found, use default setting", stacklevel=2)
    images["mask"] = np.full(images["src"].shape, 255, np.uint8)
  else:
    images["mask"] = read_image(mask_name)
  return images["src"], images["mask"], images["tgt"]
