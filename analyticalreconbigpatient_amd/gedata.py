"""GE helical projection files: the header and segment readers of the reference.

* ``GetHeadofGeDataSet`` -- src/GetHeadofGeDataSet.m:16-30: 18 float32 values;
  the fields used are (0-based) 2 SDD, 3 SO, 4 cell pitch, 5 first view angle,
  8 views/rotation, 9 total views, 10 channels/row, 11 iso channel, 12 rows,
  13 slice height, 14 helical pitch.
* ``readBigPatientData`` -- src/readBigPatientData.m:15-34: a headerless file of
  float32 projections, channel fastest ([view][row][channel]); segment SegInd
  starts (SegLength - 2 SegOverLap) * (SegInd - 1) views in and holds SegLength
  views (the last one whatever remains).  Returned as (DLength, YL, ZL), the
  MATLAB permute(Proj, [3 1 2]).

The projection file is memory-mapped, so a segment costs only its own bytes.
``write_ge_header`` / ``write_projections`` produce files of the same format
(used for synthetic scans).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .geometry import GEHeader, seg_count

_HDR_FIELDS = {  # 0-based float index -> GEHeader field (GetHeadofGeDataSet.m:20-30)
    2: "Source2DetDis", 3: "Source2IsoDis", 4: "DetCellPitch", 5: "BeginViewAngle",
    8: "NumberOfViewPerRotation", 9: "TotalViewNumber", 10: "NumberOfDetChannelPerRow",
    11: "IsoChannelLocation", 12: "NumberOfDetectorRow", 13: "SliceHeigh", 14: "HelicalPitch",
}
_INT_FIELDS = {"NumberOfViewPerRotation", "TotalViewNumber", "NumberOfDetChannelPerRow",
               "NumberOfDetectorRow"}


def GetHeadofGeDataSet(path) -> GEHeader:
    raw = np.fromfile(path, dtype="<f4", count=18)
    if raw.size != 18:
        raise ValueError(f"{path}: a GE header holds 18 float32 values, found {raw.size}")
    kw = {}
    for i, name in _HDR_FIELDS.items():
        v = float(raw[i])
        kw[name] = int(round(v)) if name in _INT_FIELDS else v
    return GEHeader(**kw)


def write_ge_header(path, hdr: GEHeader) -> None:
    raw = np.zeros(18, dtype="<f4")
    for i, name in _HDR_FIELDS.items():
        raw[i] = getattr(hdr, name)
    raw.tofile(path)


def segment_span(hdr: GEHeader, SegLength: int, SegOverLap: int, SegInd: int) -> tuple[int, int]:
    """(first view, DLength) of segment SegInd (1-based), readBigPatientData.m:16-24."""
    SegNum = seg_count(hdr, SegLength, SegOverLap)
    if SegInd < SegNum:
        DLength = SegLength
    else:
        DLength = hdr.TotalViewNumber - (SegLength - 2 * SegOverLap) * (SegInd - 1)
    first = (SegLength - 2 * SegOverLap) * (SegInd - 1)
    return first, DLength


def readBigPatientData(ProjName, CTGE: GEHeader, SegLength: int, SegOverLap: int,
                       SegInd: int):
    """Returns (Proj (DLength, YL, ZL) float32 view of the file, DLength)."""
    YL, ZL = CTGE.NumberOfDetChannelPerRow, CTGE.NumberOfDetectorRow
    first, DLength = segment_span(CTGE, SegLength, SegOverLap, SegInd)
    if first < 0 or DLength <= 0 or first + DLength > CTGE.TotalViewNumber:
        raise ValueError(f"segment {SegInd} lies outside the scan")
    mm = np.memmap(ProjName, dtype="<f4", mode="r", offset=4 * first * YL * ZL,
                   shape=(DLength, ZL, YL))          # [view][row][channel]
    return np.transpose(mm, (0, 2, 1)), DLength      # (DLength, YL, ZL)


def write_projections(path, proj_vyz: np.ndarray) -> None:
    """Writes (ViewN, YL, ZL) projections in the file order [view][row][channel]."""
    np.ascontiguousarray(np.transpose(proj_vyz, (0, 2, 1)), dtype="<f4").tofile(path)


__all__ = ["GetHeadofGeDataSet", "readBigPatientData", "write_ge_header", "write_projections",
           "segment_span"]
