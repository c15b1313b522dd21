"""A small DICOM Part 10 reader and writer for the plan layer's files: CT
series, RT Dose, RT Structure Set, RT Plan and RT Ion Plan (the files the
reference opens with pydicom in plan.py:35-49,91-163, plan_imrt.py:273-351
and plan_impt.py:298-643).

Encodings: Implicit VR Little Endian (1.2.840.10008.1.2) and Explicit VR
Little Endian (1.2.840.10008.1.2.1), with sequences and items of defined or
undefined length (PS3.5 sections 7.1, 7.5).  Compressed or big-endian
transfer syntaxes raise NotImplementedError.  Files are written as Explicit
VR Little Endian with a 128-byte preamble and a File Meta group (PS3.10).

Values follow the usual Python mapping: DS and IS become float / int
subclasses that keep their text (``DS("6") == "6"`` holds, so energy labels
compare like strings), multi-valued elements become lists, SQ becomes a list
of Dataset, OB/OW/UN stay bytes.  Attributes are reached by keyword for the
tags in ``KEYWORDS`` and by (group, element) tuple for any tag.
"""
from __future__ import annotations

import os
import struct
import uuid

import numpy as np

IMPLICIT_LE = "1.2.840.10008.1.2"
EXPLICIT_LE = "1.2.840.10008.1.2.1"
CT_IMAGE_STORAGE = "1.2.840.10008.5.1.4.1.1.2"
RT_DOSE_STORAGE = "1.2.840.10008.5.1.4.1.1.481.2"
RT_STRUCT_STORAGE = "1.2.840.10008.5.1.4.1.1.481.3"
RT_PLAN_STORAGE = "1.2.840.10008.5.1.4.1.1.481.5"
RT_ION_PLAN_STORAGE = "1.2.840.10008.5.1.4.1.1.481.8"
IMPLEMENTATION_UID = "2.25.91230517640372398817410911651962617861"

# keyword -> (group, element, VR): the attributes the plan layer reads/writes
KEYWORDS = {
    # file meta (PS3.10 7.1)
    "FileMetaInformationGroupLength": (0x0002, 0x0000, "UL"),
    "FileMetaInformationVersion": (0x0002, 0x0001, "OB"),
    "MediaStorageSOPClassUID": (0x0002, 0x0002, "UI"),
    "MediaStorageSOPInstanceUID": (0x0002, 0x0003, "UI"),
    "TransferSyntaxUID": (0x0002, 0x0010, "UI"),
    "ImplementationClassUID": (0x0002, 0x0012, "UI"),
    "ImplementationVersionName": (0x0002, 0x0013, "SH"),
    # patient / study / series / instance
    "InstanceCreationDate": (0x0008, 0x0012, "DA"),
    "InstanceCreationTime": (0x0008, 0x0013, "TM"),
    "SOPClassUID": (0x0008, 0x0016, "UI"),
    "SOPInstanceUID": (0x0008, 0x0018, "UI"),
    "StudyDate": (0x0008, 0x0020, "DA"),
    "Modality": (0x0008, 0x0060, "CS"),
    "Manufacturer": (0x0008, 0x0070, "LO"),
    "SeriesDescription": (0x0008, 0x103E, "LO"),
    "ReferencedSOPClassUID": (0x0008, 0x1150, "UI"),
    "ReferencedSOPInstanceUID": (0x0008, 0x1155, "UI"),
    "PatientName": (0x0010, 0x0010, "PN"),
    "PatientID": (0x0010, 0x0020, "LO"),
    "PatientIdentityRemoved": (0x0012, 0x0062, "CS"),
    "DeidentificationMethod": (0x0012, 0x0063, "LO"),
    "SliceThickness": (0x0018, 0x0050, "DS"),
    "PatientPosition": (0x0018, 0x5100, "CS"),
    "StudyInstanceUID": (0x0020, 0x000D, "UI"),
    "SeriesInstanceUID": (0x0020, 0x000E, "UI"),
    "SeriesNumber": (0x0020, 0x0011, "IS"),
    "InstanceNumber": (0x0020, 0x0013, "IS"),
    "ImagePositionPatient": (0x0020, 0x0032, "DS"),
    "ImageOrientationPatient": (0x0020, 0x0037, "DS"),
    "FrameOfReferenceUID": (0x0020, 0x0052, "UI"),
    # image pixel
    "SamplesPerPixel": (0x0028, 0x0002, "US"),
    "PhotometricInterpretation": (0x0028, 0x0004, "CS"),
    "NumberOfFrames": (0x0028, 0x0008, "IS"),
    "FrameIncrementPointer": (0x0028, 0x0009, "AT"),
    "Rows": (0x0028, 0x0010, "US"),
    "Columns": (0x0028, 0x0011, "US"),
    "PixelSpacing": (0x0028, 0x0030, "DS"),
    "BitsAllocated": (0x0028, 0x0100, "US"),
    "BitsStored": (0x0028, 0x0101, "US"),
    "HighBit": (0x0028, 0x0102, "US"),
    "PixelRepresentation": (0x0028, 0x0103, "US"),
    "RescaleIntercept": (0x0028, 0x1052, "DS"),
    "RescaleSlope": (0x0028, 0x1053, "DS"),
    "RescaleType": (0x0028, 0x1054, "LO"),
    "PixelData": (0x7FE0, 0x0010, "OW"),
    # RT dose
    "DoseUnits": (0x3004, 0x0002, "CS"),
    "DoseType": (0x3004, 0x0004, "CS"),
    "DoseSummationType": (0x3004, 0x000A, "CS"),
    "GridFrameOffsetVector": (0x3004, 0x000C, "DS"),
    "DoseGridScaling": (0x3004, 0x000E, "DS"),
    # RT structure set
    "StructureSetROISequence": (0x3006, 0x0020, "SQ"),
    "ROINumber": (0x3006, 0x0022, "IS"),
    "ROIName": (0x3006, 0x0026, "LO"),
    "ReferencedROINumber": (0x3006, 0x0084, "IS"),
    # RT (ion) plan
    "RTPlanLabel": (0x300A, 0x0002, "SH"),
    "RTPlanDate": (0x300A, 0x0006, "DA"),
    "RTPlanTime": (0x300A, 0x0007, "TM"),
    "TreatmentProtocols": (0x300A, 0x0009, "LO"),
    "PlanIntent": (0x300A, 0x000A, "CS"),
    "RTPlanGeometry": (0x300A, 0x000C, "CS"),
    "DoseReferenceSequence": (0x300A, 0x0010, "SQ"),
    "DoseReferenceNumber": (0x300A, 0x0012, "IS"),
    "DoseReferenceStructureType": (0x300A, 0x0014, "CS"),
    "NominalBeamEnergyUnit": (0x300A, 0x0015, "CS"),
    "DoseReferenceDescription": (0x300A, 0x0016, "LO"),
    "DoseReferenceType": (0x300A, 0x0020, "CS"),
    "TargetPrescriptionDose": (0x300A, 0x0026, "DS"),
    "TargetUnderdoseVolumeFraction": (0x300A, 0x0028, "DS"),
    "FractionGroupSequence": (0x300A, 0x0070, "SQ"),
    "FractionGroupNumber": (0x300A, 0x0071, "IS"),
    "NumberOfFractionsPlanned": (0x300A, 0x0078, "IS"),
    "NumberOfBeams": (0x300A, 0x0080, "IS"),
    "BeamMeterset": (0x300A, 0x0086, "DS"),
    "NumberOfBrachyApplicationSetups": (0x300A, 0x00A0, "IS"),
    "BeamSequence": (0x300A, 0x00B0, "SQ"),
    "TreatmentMachineName": (0x300A, 0x00B2, "SH"),
    "PrimaryDosimeterUnit": (0x300A, 0x00B3, "CS"),
    "RTBeamLimitingDeviceType": (0x300A, 0x00B8, "CS"),
    "BeamNumber": (0x300A, 0x00C0, "IS"),
    "BeamName": (0x300A, 0x00C2, "LO"),
    "BeamType": (0x300A, 0x00C4, "CS"),
    "RadiationType": (0x300A, 0x00C6, "CS"),
    "TreatmentDeliveryType": (0x300A, 0x00CE, "CS"),
    "NumberOfWedges": (0x300A, 0x00D0, "IS"),
    "NumberOfCompensators": (0x300A, 0x00E0, "IS"),
    "NumberOfBoli": (0x300A, 0x00ED, "IS"),
    "NumberOfBlocks": (0x300A, 0x00F0, "IS"),
    "AccessoryCode": (0x300A, 0x00F9, "LO"),
    "CumulativeDoseReferenceCoefficient": (0x300A, 0x010C, "DS"),
    "FinalCumulativeMetersetWeight": (0x300A, 0x010E, "DS"),
    "NumberOfControlPoints": (0x300A, 0x0110, "IS"),
    "ControlPointSequence": (0x300A, 0x0111, "SQ"),
    "ControlPointIndex": (0x300A, 0x0112, "IS"),
    "NominalBeamEnergy": (0x300A, 0x0114, "DS"),
    "BeamLimitingDevicePositionSequence": (0x300A, 0x011A, "SQ"),
    "LeafJawPositions": (0x300A, 0x011C, "DS"),
    "GantryAngle": (0x300A, 0x011E, "DS"),
    "GantryRotationDirection": (0x300A, 0x011F, "CS"),
    "BeamLimitingDeviceAngle": (0x300A, 0x0120, "DS"),
    "BeamLimitingDeviceRotationDirection": (0x300A, 0x0121, "CS"),
    "PatientSupportAngle": (0x300A, 0x0122, "DS"),
    "PatientSupportRotationDirection": (0x300A, 0x0123, "CS"),
    "TableTopEccentricAngle": (0x300A, 0x0125, "DS"),
    "TableTopEccentricRotationDirection": (0x300A, 0x0126, "CS"),
    "TableTopVerticalPosition": (0x300A, 0x0128, "DS"),
    "TableTopLongitudinalPosition": (0x300A, 0x0129, "DS"),
    "TableTopLateralPosition": (0x300A, 0x012A, "DS"),
    "IsocenterPosition": (0x300A, 0x012C, "DS"),
    "CumulativeMetersetWeight": (0x300A, 0x0134, "DS"),
    "TableTopPitchAngle": (0x300A, 0x0140, "FL"),
    "TableTopPitchRotationDirection": (0x300A, 0x0142, "CS"),
    "TableTopRollAngle": (0x300A, 0x0144, "FL"),
    "TableTopRollRotationDirection": (0x300A, 0x0146, "CS"),
    "GantryPitchAngle": (0x300A, 0x014A, "FL"),
    "GantryPitchRotationDirection": (0x300A, 0x014C, "CS"),
    "PatientSetupSequence": (0x300A, 0x0180, "SQ"),
    "PatientSetupNumber": (0x300A, 0x0182, "IS"),
    "ScanMode": (0x300A, 0x0308, "CS"),
    "ModulatedScanModeType": (0x300A, 0x0309, "CS"),
    "VirtualSourceAxisDistances": (0x300A, 0x030A, "FL"),
    "SnoutSequence": (0x300A, 0x030C, "SQ"),
    "SnoutPosition": (0x300A, 0x030D, "FL"),
    "SnoutID": (0x300A, 0x030F, "SH"),
    "NumberOfRangeShifters": (0x300A, 0x0312, "IS"),
    "RangeShifterSequence": (0x300A, 0x0314, "SQ"),
    "RangeShifterNumber": (0x300A, 0x0316, "IS"),
    "RangeShifterID": (0x300A, 0x0318, "SH"),
    "RangeShifterType": (0x300A, 0x0320, "CS"),
    "NumberOfLateralSpreadingDevices": (0x300A, 0x0330, "IS"),
    "LateralSpreadingDeviceSequence": (0x300A, 0x0332, "SQ"),
    "LateralSpreadingDeviceNumber": (0x300A, 0x0334, "IS"),
    "LateralSpreadingDeviceID": (0x300A, 0x0336, "SH"),
    "LateralSpreadingDeviceType": (0x300A, 0x0338, "CS"),
    "NumberOfRangeModulators": (0x300A, 0x0340, "IS"),
    "RangeModulatorSequence": (0x300A, 0x0342, "SQ"),
    "RangeModulatorNumber": (0x300A, 0x0344, "IS"),
    "RangeModulatorID": (0x300A, 0x0346, "SH"),
    "RangeModulatorType": (0x300A, 0x0348, "CS"),
    "RangeModulatorDescription": (0x300A, 0x034A, "LO"),
    "PatientSupportType": (0x300A, 0x0350, "CS"),
    "PatientSupportID": (0x300A, 0x0352, "SH"),
    "MetersetRate": (0x300A, 0x035A, "FL"),
    "RangeShifterSetting": (0x300A, 0x0362, "LO"),
    "LateralSpreadingDeviceSettingsSequence": (0x300A, 0x0370, "SQ"),
    "LateralSpreadingDeviceSetting": (0x300A, 0x0372, "LO"),
    "IsocenterToLateralSpreadingDeviceDistance": (0x300A, 0x0374, "FL"),
    "LateralSpreadingDeviceWaterEquivalentThickness": (0x300A, 0x037C, "FL"),
    "RangeModulatorSettingsSequence": (0x300A, 0x0380, "SQ"),
    "ScanSpotTuneID": (0x300A, 0x0390, "SH"),
    "NumberOfScanSpotPositions": (0x300A, 0x0392, "IS"),
    "ScanSpotPositionMap": (0x300A, 0x0394, "FL"),
    "ScanSpotMetersetWeights": (0x300A, 0x0396, "FL"),
    "ScanningSpotSize": (0x300A, 0x0398, "FL"),
    "NumberOfPaintings": (0x300A, 0x039A, "IS"),
    "IonBeamSequence": (0x300A, 0x03A2, "SQ"),
    "IonControlPointSequence": (0x300A, 0x03A8, "SQ"),
    "ReferencedRTPlanSequence": (0x300C, 0x0002, "SQ"),
    "ReferencedBeamSequence": (0x300C, 0x0004, "SQ"),
    "ReferencedBeamNumber": (0x300C, 0x0006, "IS"),
    "ReferencedDoseReferenceSequence": (0x300C, 0x0050, "SQ"),
    "ReferencedDoseReferenceNumber": (0x300C, 0x0051, "IS"),
    "ReferencedStructureSetSequence": (0x300C, 0x0060, "SQ"),
    "ReferencedRangeShifterNumber": (0x300C, 0x0100, "IS"),
    "ReferencedLateralSpreadingDeviceNumber": (0x300C, 0x0102, "IS"),
    "ReferencedRangeModulatorNumber": (0x300C, 0x0104, "IS"),
    "ApprovalStatus": (0x300E, 0x0002, "CS"),
}
_BY_TAG = {(g, e): (kw, vr) for kw, (g, e, vr) in KEYWORDS.items()}

# VRs with a 2-byte reserved field and a 4-byte length in explicit VR (PS3.5 7.1.2)
_LONG_VRS = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR", "UT", "UV"}
_TEXT_VRS = {"AE", "AS", "CS", "DA", "DT", "LO", "LT", "PN", "SH", "ST", "TM", "UC", "UI", "UR", "UT"}
_SINGLE_TEXT = {"LT", "ST", "UT", "UR"}      # no backslash splitting
_BINARY = {"US": "<H", "SS": "<h", "UL": "<I", "SL": "<i", "FL": "<f", "FD": "<d",
           "SV": "<q", "UV": "<Q"}
_ITEM, _ITEM_END, _SEQ_END = (0xFFFE, 0xE000), (0xFFFE, 0xE00D), (0xFFFE, 0xE0DD)
_UNDEFINED = 0xFFFFFFFF


class DS(float):
    """Decimal String value: a float that keeps its text."""

    def __new__(cls, text):
        if isinstance(text, str):
            t = text.strip()
            obj = float.__new__(cls, float(t))
            obj.text = t
        else:
            obj = float.__new__(cls, float(text))
            obj.text = _format_ds(float(text))
        return obj

    def __str__(self):
        return self.text

    __repr__ = __str__

    def __eq__(self, other):
        if isinstance(other, str):
            return self.text == other.strip()
        return float(self) == other

    def __ne__(self, other):
        return not self.__eq__(other)

    __hash__ = float.__hash__


class IS(int):
    """Integer String value: an int that keeps its text."""

    def __new__(cls, text):
        if isinstance(text, str):
            t = text.strip()
            obj = int.__new__(cls, int(float(t)) if t else 0)
            obj.text = t
        else:
            obj = int.__new__(cls, int(text))
            obj.text = str(int(text))
        return obj

    def __str__(self):
        return self.text

    __repr__ = __str__

    def __eq__(self, other):
        if isinstance(other, str):
            return self.text == other.strip()
        return int(self) == other

    def __ne__(self, other):
        return not self.__eq__(other)

    __hash__ = int.__hash__


def _format_ds(x: float) -> str:
    """A Decimal String of at most 16 characters (PS3.5 6.2) that reads back
    as the closest value it can hold."""
    if x == int(x) and abs(x) < 1e15:
        return str(int(x))
    for prec in range(16, 0, -1):
        s = "%.*g" % (prec, x)
        if "e" in s:
            m, e = s.split("e")
            s = m + "e" + str(int(e))
        if len(s) <= 16:
            return s
    raise ValueError("value %r does not fit a Decimal String" % x)


class Element:
    __slots__ = ("tag", "VR", "value")

    def __init__(self, tag, VR, value):
        self.tag, self.VR, self.value = (int(tag[0]), int(tag[1])), VR, value

    def __repr__(self):
        return "(%04X,%04X) %s %r" % (self.tag[0], self.tag[1], self.VR, self.value)


DataElement = Element


class Dataset:
    """An ordered set of elements, reached by keyword or (group, element)."""

    def __init__(self):
        object.__setattr__(self, "_el", {})

    # -- element access
    def __getitem__(self, tag):
        return self._el[(int(tag[0]), int(tag[1]))]

    def __setitem__(self, tag, el):
        if not isinstance(el, Element):
            raise TypeError("Dataset items are Element objects")
        self._el[(int(tag[0]), int(tag[1]))] = el

    def __delitem__(self, tag):
        del self._el[(int(tag[0]), int(tag[1]))]

    def __contains__(self, key):
        if isinstance(key, str):
            return key in KEYWORDS and KEYWORDS[key][:2] in self._el
        return (int(key[0]), int(key[1])) in self._el

    def __iter__(self):
        return iter(self._el[t] for t in sorted(self._el))

    def __len__(self):
        return len(self._el)

    def add_new(self, tag, VR, value):
        self[tag] = Element(tag, VR, value)

    # -- keyword access
    def __getattr__(self, name):
        kw = KEYWORDS.get(name)
        if kw is None or kw[:2] not in self._el:
            raise AttributeError(name)
        return self._el[kw[:2]].value

    def __setattr__(self, name, value):
        kw = KEYWORDS.get(name)
        if kw is None:
            raise AttributeError("'%s' is not a DICOM keyword this module knows" % name)
        g, e, vr = kw
        self._el[(g, e)] = Element((g, e), vr, _coerce(vr, value))

    def __delattr__(self, name):
        kw = KEYWORDS.get(name)
        if kw is None or kw[:2] not in self._el:
            raise AttributeError(name)
        del self._el[kw[:2]]

    def get(self, name, default=None):
        try:
            return getattr(self, name)
        except AttributeError:
            return default

    # -- pixels
    @property
    def pixel_array(self) -> np.ndarray:
        """The pixel data as (frames, rows, columns) or (rows, columns)."""
        bits = int(self.BitsAllocated)
        signed = int(self.get("PixelRepresentation", 0)) == 1
        if int(self.get("SamplesPerPixel", 1)) != 1:
            raise NotImplementedError("only single-sample (grey-scale) pixel data is read")
        dt = {(8, False): "<u1", (8, True): "<i1", (16, False): "<u2", (16, True): "<i2",
              (32, False): "<u4", (32, True): "<i4"}.get((bits, signed))
        if dt is None:
            raise NotImplementedError("BitsAllocated %d is not read" % bits)
        rows, cols = int(self.Rows), int(self.Columns)
        frames = int(self.get("NumberOfFrames", 1) or 1)
        n = rows * cols * frames
        a = np.frombuffer(bytes(self.PixelData), dtype=dt)
        if a.size < n:
            raise ValueError("PixelData holds %d values, the header asks for %d" % (a.size, n))
        a = a[:n]
        return a.reshape(frames, rows, cols) if frames > 1 else a.reshape(rows, cols)


def _coerce(vr, value):
    """Python value for an element of this VR (what the reader would give)."""
    if value is None:
        return None
    if vr == "SQ":
        return list(value)
    if vr in ("DS", "IS"):
        cls = DS if vr == "DS" else IS
        if isinstance(value, str):
            if value.strip() == "":
                return None
            parts = value.split("\\")
            return cls(parts[0]) if len(parts) == 1 else [cls(p) for p in parts]
        if isinstance(value, (list, tuple, np.ndarray)):
            a = np.asarray(value)
            return [cls(v) for v in a.ravel().tolist()] if a.ndim else cls(a.item())
        return cls(value)
    if vr in _BINARY:
        conv = float if vr in ("FL", "FD") else int
        if isinstance(value, (list, tuple, np.ndarray)) and np.ndim(value):
            return [conv(v) for v in np.asarray(value).ravel().tolist()]
        return conv(value)
    if vr in ("OB", "OW", "OF", "OD", "OL", "UN"):
        if isinstance(value, np.ndarray):
            return value.tobytes()
        return bytes(value) if not isinstance(value, bytes) else value
    if vr == "AT":
        return value
    if isinstance(value, (list, tuple)):
        return [str(v) for v in value]
    return str(value)


# ---------------------------------------------------------------------------
# decoding
# ---------------------------------------------------------------------------
def _decode_value(vr, raw: bytes):
    if vr in _TEXT_VRS or vr in ("DS", "IS"):
        s = raw.decode("latin-1").rstrip("\x00 ")
        if vr == "UI":
            s = s.rstrip("\x00")
        if vr in ("DS", "IS"):
            if s.strip() == "":
                return None
            parts = s.split("\\")
            cls = DS if vr == "DS" else IS
            return cls(parts[0]) if len(parts) == 1 else [cls(p) for p in parts]
        if vr in _SINGLE_TEXT:
            return s
        parts = s.split("\\")
        return parts[0] if len(parts) == 1 else parts
    if vr in _BINARY:
        fmt = _BINARY[vr]
        size = struct.calcsize(fmt)
        if len(raw) == 0:
            return None
        vals = list(struct.unpack("<%d%s" % (len(raw) // size, fmt[1]), raw[:len(raw) // size * size]))
        return vals[0] if len(vals) == 1 else vals
    if vr == "AT":
        vals = struct.unpack("<%dH" % (len(raw) // 2), raw)
        pairs = [(vals[i], vals[i + 1]) for i in range(0, len(vals), 2)]
        return pairs[0] if len(pairs) == 1 else pairs
    return raw


class _Reader:
    def __init__(self, buf: bytes, explicit: bool):
        self.b, self.explicit = buf, explicit

    def tag(self, p):
        g, e = struct.unpack_from("<HH", self.b, p)
        return (g, e)

    def dataset(self, p, end, stop_at_item_end=False):
        """Elements from p up to end (or up to an item delimiter); returns
        (Dataset, position after the last byte read)."""
        ds = Dataset()
        while p < end:
            t = self.tag(p)
            if t == _ITEM_END:
                if not stop_at_item_end:
                    raise ValueError("unexpected item delimiter at offset %d" % p)
                return ds, p + 8
            el, p = self.element(p)
            ds._el[el.tag] = el
        return ds, p

    def element(self, p):
        t = self.tag(p)
        if self.explicit and t[0] != 0xFFFE:
            vr = self.b[p + 4:p + 6].decode("latin-1")
            if vr in _LONG_VRS:
                (length,) = struct.unpack_from("<I", self.b, p + 8)
                p += 12
            elif vr.isalpha() and vr.isupper():
                (length,) = struct.unpack_from("<H", self.b, p + 6)
                p += 8
            else:
                raise ValueError("bad VR %r at offset %d" % (vr, p))
        else:
            vr = _BY_TAG.get(t, (None, "UN"))[1]
            if t[1] == 0x0000:
                vr = "UL"                      # group lengths
            (length,) = struct.unpack_from("<I", self.b, p + 4)
            p += 8
        if length == _UNDEFINED and t == (0x7FE0, 0x0010):
            raise NotImplementedError("encapsulated (compressed) pixel data is not read")
        if vr == "SQ":
            items, p = self.sequence(p, length)
            return Element(t, "SQ", items), p
        if length == _UNDEFINED and vr == "UN":
            # an unknown-VR sequence: its items are implicit VR (PS3.5 6.2.2)
            items, p = _Reader(self.b, explicit=False).sequence(p, length)
            return Element(t, "SQ", items), p
        if length == _UNDEFINED:
            raise ValueError("undefined length on a %s element at offset %d" % (vr, p))
        raw = self.b[p:p + length]
        if len(raw) != length:
            raise ValueError("element (%04X,%04X) runs past the end of the file" % t)
        return Element(t, vr, _decode_value(vr, raw)), p + length

    def sequence(self, p, length):
        items = []
        end = len(self.b) if length == _UNDEFINED else p + length
        while p < end:
            t = self.tag(p)
            (ilen,) = struct.unpack_from("<I", self.b, p + 4)
            p += 8
            if t == _SEQ_END:
                return items, p
            if t != _ITEM:
                raise ValueError("expected an item tag at offset %d, found (%04X,%04X)" % (p - 8, *t))
            if ilen == _UNDEFINED:
                ds, p = self.dataset(p, len(self.b), stop_at_item_end=True)
            else:
                ds, _ = self.dataset(p, p + ilen)
                p += ilen
            items.append(ds)
        if length == _UNDEFINED:
            raise ValueError("sequence without a delimiter")
        return items, p


def read(path_or_bytes, force: bool = False) -> Dataset:
    """A DICOM file (path or bytes).  With the 'DICM' prefix the File Meta
    group names the transfer syntax; without it (force=True, as pydicom's
    dcmread(force=True)) the encoding is recognised from the first element."""
    buf = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    buf = bytes(buf)
    p = 0
    meta = Dataset()
    if len(buf) >= 132 and buf[128:132] == b"DICM":
        p = 132
    elif buf[:4] == b"DICM":
        p = 4
    elif not force:
        raise ValueError("not a DICOM Part 10 file (no 'DICM' prefix); use force=True")
    if p:
        r = _Reader(buf, explicit=True)
        while p < len(buf) and r.tag(p)[0] == 0x0002:
            el, p = r.element(p)
            meta._el[el.tag] = el
        ts = str(meta.get("TransferSyntaxUID", EXPLICIT_LE)).strip("\x00 ")
    else:
        vr = buf[4:6]
        ts = EXPLICIT_LE if len(buf) >= 6 and vr.isalpha() and vr.isupper() else IMPLICIT_LE
    if ts not in (IMPLICIT_LE, EXPLICIT_LE):
        raise NotImplementedError("transfer syntax %s is not read (only uncompressed little endian)" % ts)
    ds, _ = _Reader(buf, explicit=(ts == EXPLICIT_LE)).dataset(p, len(buf))
    object.__setattr__(ds, "file_meta", meta)
    return ds


dcmread = read


# ---------------------------------------------------------------------------
# encoding (explicit VR little endian)
# ---------------------------------------------------------------------------
def _encode_value(vr, value) -> bytes:
    if value is None:
        return b""
    if vr in ("DS", "IS"):
        vals = value if isinstance(value, list) else [value]
        s = "\\".join(v.text if hasattr(v, "text") else (_format_ds(float(v)) if vr == "DS" else str(int(v)))
                      for v in vals)
        raw = s.encode("latin-1")
        return raw + b" " if len(raw) % 2 else raw
    if vr in _TEXT_VRS:
        s = "\\".join(value) if isinstance(value, list) else str(value)
        raw = s.encode("latin-1")
        if len(raw) % 2:
            raw += b"\x00" if vr == "UI" else b" "
        return raw
    if vr in _BINARY:
        vals = value if isinstance(value, list) else [value]
        return struct.pack("<%d%s" % (len(vals), _BINARY[vr][1]), *vals)
    if vr == "AT":
        pairs = value if isinstance(value, list) else [value]
        return b"".join(struct.pack("<HH", *p) for p in pairs)
    raw = bytes(value)
    return raw + b"\x00" if len(raw) % 2 else raw


def _encode_element(el: Element) -> bytes:
    g, e = el.tag
    vr = el.VR
    if vr == "SQ":
        body = b""
        for item in (el.value or []):
            ib = _encode_dataset(item)
            body += struct.pack("<HHI", 0xFFFE, 0xE000, len(ib)) + ib
        return struct.pack("<HH", g, e) + b"SQ\x00\x00" + struct.pack("<I", len(body)) + body
    raw = _encode_value(vr, el.value)
    if vr in _LONG_VRS:
        return struct.pack("<HH", g, e) + vr.encode() + b"\x00\x00" + struct.pack("<I", len(raw)) + raw
    if len(raw) > 0xFFFF:
        raise ValueError("(%04X,%04X) %s value of %d bytes exceeds its 16-bit length" % (g, e, vr, len(raw)))
    return struct.pack("<HH", g, e) + vr.encode() + struct.pack("<H", len(raw)) + raw


def _encode_dataset(ds: Dataset) -> bytes:
    return b"".join(_encode_element(el) for el in ds if el.tag[0] != 0x0002)


def generate_uid() -> str:
    """A UUID-derived UID under the 2.25 root (PS3.5 B.2)."""
    return "2.25.%d" % uuid.uuid4().int


def write(ds: Dataset, path: str, sop_class_uid: str = None, sop_instance_uid: str = None):
    """The dataset as an Explicit VR Little Endian Part 10 file."""
    cls = sop_class_uid or ds.get("SOPClassUID")
    inst = sop_instance_uid or ds.get("SOPInstanceUID")
    if cls is None or inst is None:
        raise ValueError("writing a DICOM file needs its SOPClassUID and SOPInstanceUID")
    meta = Dataset()
    meta.FileMetaInformationVersion = b"\x00\x01"
    meta.MediaStorageSOPClassUID = str(cls)
    meta.MediaStorageSOPInstanceUID = str(inst)
    meta.TransferSyntaxUID = EXPLICIT_LE
    meta.ImplementationClassUID = IMPLEMENTATION_UID
    meta.ImplementationVersionName = "DOSEHIP_1"
    body = b"".join(_encode_element(el) for el in meta)
    head = _encode_element(Element((0x0002, 0x0000), "UL", len(body)))
    with open(path, "wb") as f:
        f.write(b"\x00" * 128 + b"DICM" + head + body + _encode_dataset(ds))


def save_as(ds: Dataset, path: str):
    write(ds, path)


def list_files(directory: str, suffix: str = None):
    names = sorted(os.listdir(directory))
    return [os.path.join(directory, n) for n in names
            if os.path.isfile(os.path.join(directory, n)) and (suffix is None or n.lower().endswith(suffix))]
