"""Image I/O (fpie/io.py:10-39) and the command line (fpie/cli.py:10-63).

CPU: Pillow-backed read/write keep cv2's BGR order, gray images become three
channels, alpha is dropped, a missing mask file gives the all-255 mask with
a warning.  GPU: the command line on the hip backend writes the image the
processor computes, with the -p snapshots, for both methods.
"""

import os
import sys

import numpy as np
import pytest

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))), "fast-poisson-image-editing_amd")
if PKG not in sys.path:
  sys.path.insert(0, PKG)


def test_write_then_read_is_identity_in_bgr(tmp_path):
  from PIL import Image
  from fpie_amd import io
  rng = np.random.default_rng(0)
  img = rng.integers(0, 256, (13, 17, 3), dtype=np.uint8)
  name = str(tmp_path / "a.png")
  io.write_image(name, img)
  np.testing.assert_array_equal(io.read_image(name), img)
  # on disk the channels are RGB: the array's last channel comes first
  disk = np.asarray(Image.open(name))
  np.testing.assert_array_equal(disk, img[..., ::-1])


def test_gray_and_alpha_images(tmp_path):
  from PIL import Image
  from fpie_amd import io
  gray = np.arange(20, dtype=np.uint8).reshape(4, 5)
  Image.fromarray(gray).save(tmp_path / "g.png")
  g = io.read_image(str(tmp_path / "g.png"))
  assert g.shape == (4, 5, 3)
  for c in range(3):
    np.testing.assert_array_equal(g[..., c], gray)
  rgba = np.zeros((4, 5, 4), np.uint8)
  rgba[..., 0], rgba[..., 3] = 200, 7
  Image.fromarray(rgba).save(tmp_path / "a.png")
  a = io.read_image(str(tmp_path / "a.png"))
  assert a.shape == (4, 5, 3)
  assert (a[..., 2] == 200).all() and (a[..., :2] == 0).all()


def test_missing_mask_is_all_255(tmp_path):
  from fpie_amd import io
  img = np.full((6, 7, 3), 9, np.uint8)
  io.write_image(str(tmp_path / "s.png"), img)
  io.write_image(str(tmp_path / "t.png"), img)
  with pytest.warns(UserWarning, match="No mask"):
    src, mask, tgt = io.read_images(str(tmp_path / "s.png"),
                                    str(tmp_path / "none.png"),
                                    str(tmp_path / "t.png"))
  assert mask.shape == src.shape and (mask == 255).all()
  with pytest.raises(FileNotFoundError):
    io.read_image(str(tmp_path / "none.png"))


def _scene(tmp_path):
  from fpie_amd import io
  rng = np.random.default_rng(5)
  src = rng.integers(0, 256, (40, 50, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (60, 70, 3), dtype=np.uint8)
  mask = np.zeros((40, 50, 3), np.uint8)
  mask[8:30, 10:42] = 255
  for name, a in (("src", src), ("tgt", tgt), ("mask", mask)):
    io.write_image(str(tmp_path / f"{name}.png"), a)
  return src, mask, tgt


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["equ", "grid"])
def test_cli_writes_the_processor_image(tmp_path, monkeypatch, method):
  from fpie_amd import cli, io, process
  src, mask, tgt = _scene(tmp_path)
  monkeypatch.chdir(tmp_path)
  cli.main(["-s", "src.png", "-m", "mask.png", "-t", "tgt.png",
            "-o", "out.png", "-b", "hip", "--method", method, "-g", "avg",
            "-n", "60", "-p", "20", "-h1", "5", "-w1", "9"])
  cls = process.EquProcessor if method == "equ" else process.GridProcessor
  proc = cls("avg", "hip")
  proc.reset(src, mask, tgt, (0, 0), (5, 9))
  for _ in range(3):
    want, _ = proc.step(20)
  np.testing.assert_array_equal(io.read_image("out.png"), want)
  # snapshots after 20 and 40 iterations, none after the last step
  assert os.path.exists("iter00020.png") and os.path.exists("iter00040.png")
  assert not os.path.exists("iter00060.png")
