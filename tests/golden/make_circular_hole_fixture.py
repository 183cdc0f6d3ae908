"""Write circular_hole_000_001.npz from the data files the reference ships.

Reads (text/data only, with loaders that execute nothing from the files):
  img_syn_circular_hole/Circular_hole 000.tif, ... 001.tif        (8-bit frames)
  img_syn_circular_hole_maskfiles/img_syn_circular_hole_maskfiles0000.tif, ...0001.tif
  results_Circular_hole 040_ws32_st16.mat                          (scipy.io.loadmat)

and stores, in the reference's array orientation (MATLAB arrays after the
transpose on load, func_quadtree/ReadImageQuadtree.m:78 and ReadImageMasks.m:78):
  img0, img1      uint8 (H, W)      frames 000 and 001
  mask0, mask1    uint8 (H, W)      logical masks of the two frames
  ImgRefMask      uint8 (H, W)      DICpara.ImgRefMask as stored in the .mat
  gridx, gridy    int64 (2,)        DICpara.gridxyROIRange
  coordinatesFEM  float64 (N, 2)    DICmesh.coordinatesFEM (N = 1434)
  elementsFEM     int32 (E, 8)      DICmesh.elementsFEM (E = 938; columns 5-8 hanging mid-nodes)
  dirichlet       int32 (D,)        DICmesh.dirichlet
  markCoordHoleEdge int32 (K,)      DICmesh.markCoordHoleEdge
  winsize_List    float64 (N, 2)    DICpara.winsize_List
  winsize, winstepsize              DICpara scalars
  U1              float64 (2N,)     ResultDisp{1}.U  (frame 001 vs 000, after ADMM)
  F1              float64 (4N,)     ResultDefGrad{1}.F

Run once in a checkout where the reference tree is present:
  python tests/golden/make_circular_hole_fixture.py /path/to/STAQ-DIC
"""
import os
import sys

import numpy as np
import scipy.io as sio
from PIL import Image


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    imgs = [np.array(Image.open(os.path.join(ref, "img_syn_circular_hole", f"Circular_hole 00{k}.tif")))
            for k in (0, 1)]
    masks = [np.array(Image.open(os.path.join(ref, "img_syn_circular_hole_maskfiles",
                                              f"img_syn_circular_hole_maskfiles000{k}.tif")))
             for k in (0, 1)]
    for a in imgs + masks:
        assert a.ndim == 2 and a.dtype == np.uint8
    m = sio.loadmat(os.path.join(ref, "results_Circular_hole 040_ws32_st16.mat"))
    P = m["DICpara"][0, 0]
    M = m["DICmesh"][0, 0]
    roi = P["gridxyROIRange"][0, 0]
    rd = m["ResultDisp"][0, 0][0, 0]
    rf = m["ResultDefGrad"][0, 0][0, 0]
    out = dict(
        img0=imgs[0].T.copy(), img1=imgs[1].T.copy(),
        mask0=(masks[0] != 0).T.astype(np.uint8), mask1=(masks[1] != 0).T.astype(np.uint8),
        ImgRefMask=(P["ImgRefMask"] != 0).astype(np.uint8),
        gridx=np.asarray(roi["gridx"], dtype=np.int64).ravel(),
        gridy=np.asarray(roi["gridy"], dtype=np.int64).ravel(),
        coordinatesFEM=np.asarray(M["coordinatesFEM"], dtype=np.float64),
        elementsFEM=np.asarray(M["elementsFEM"], dtype=np.int32),
        dirichlet=np.asarray(M["dirichlet"], dtype=np.int32).ravel(),
        markCoordHoleEdge=np.asarray(M["markCoordHoleEdge"], dtype=np.int32).ravel(),
        winsize_List=np.asarray(P["winsize_List"], dtype=np.float64),
        winsize=np.float64(P["winsize"][0, 0]), winstepsize=np.float64(P["winstepsize"][0, 0]),
        U1=np.asarray(rd["U"], dtype=np.float64).ravel(),
        F1=np.asarray(rf["F"], dtype=np.float64).ravel(),
    )
    np.savez_compressed(os.path.join(here, "circular_hole_000_001.npz"), **out)
    for k, v in out.items():
        print(k, np.shape(v), np.asarray(v).dtype)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
