"""Scene configuration types, mirroring openPSTD's kernel API.

PSTDSettings, DomainConfEdge, DomainConf and PSTDConfiguration keep the
reference's names, getters/setters and defaults (kernel/KernelInterface.h:44-241,
kernel/KernelInterface.cpp:16-269), so code written against openPSTD's kernel
API reads the same.  ``to_c()`` produces the C structure of
include/openpstd_amd.h.
"""
import ctypes as C
import math

from . import abi_types as T

# kernel/core/kernel_functions.h:102-103
RK_COEFFICIENTS = [1.179799016570605e-1, 1.846469664911166e-1, 2.466236043095944e-1,
                   3.318395427360000e-1, 5e-1, 1.0]

PSTD_DOMAIN_SIDE_TOP = 1
PSTD_DOMAIN_SIDE_BOTTOM = 2
PSTD_DOMAIN_SIDE_LEFT = 4
PSTD_DOMAIN_SIDE_RIGHT = 8
PSTD_DOMAIN_SIDE_ALL = 15
PSTD_DOMAIN_SIDE_NONE = 0


def _f32(x):
    return float(C.c_float(x).value)


class PSTDSettings:
    """kernel/KernelInterface.h:44-158."""

    def __init__(self):
        self.calctime = 0.0
        self.c1 = 0.0
        self.ampMax = 0.0
        self.freqMax = 0.0
        self.rho = 0.0
        self.patcherror = 0.0
        self.tfactRK = 0.0
        self.gridSpacing = 0.0
        self.band_width = 0.0
        self.wave_length = 0.0
        self.rk_coefficients = list(RK_COEFFICIENTS)
        self.spectral_interpolation = False
        self.PMLCells = 0
        self.SaveNth = 1

    # getters / setters (kernel/KernelInterface.cpp:16-142)
    def GetGridSpacing(self): return self.gridSpacing
    def SetGridSpacing(self, v): self.gridSpacing = float(v)
    def GetPatchError(self): return self.patcherror
    def SetPatchError(self, v): self.patcherror = float(v)
    def GetRenderTime(self): return self.calctime
    def SetRenderTime(self, v): self.calctime = float(v)
    def GetPMLCells(self): return self.PMLCells
    def SetPMLCells(self, v): self.PMLCells = int(v)
    def GetAttenuationOfPMLCells(self): return self.ampMax
    def SetAttenuationOfPMLCells(self, v): self.ampMax = float(v)
    def GetDensityOfAir(self): return self.rho
    def SetDensityOfAir(self, v): self.rho = float(v)
    def GetMaxFrequency(self): return self.freqMax
    def SetMaxFrequency(self, v): self.freqMax = float(v)
    def GetSoundSpeed(self): return self.c1
    def SetSoundSpeed(self, v): self.c1 = float(v)
    def GetFactRK(self): return self.tfactRK
    def SetFactRK(self, v): self.tfactRK = float(v)
    def GetSaveNth(self): return self.SaveNth
    def SetSaveNth(self, v): self.SaveNth = int(v)
    def GetBandWidth(self): return self.band_width
    def SetBandWidth(self, v): self.band_width = float(v)
    def GetSpectralInterpolation(self): return self.spectral_interpolation
    def SetSpectralInterpolation(self, v): self.spectral_interpolation = bool(v)
    def GetWaveLength(self): return self.wave_length
    def SetWaveLength(self, v): self.wave_length = float(v)
    def GetRKCoefficients(self): return list(self.rk_coefficients)
    def SetRKCoefficients(self, c): self.rk_coefficients = [float(x) for x in c]

    def GetTimeStep(self):
        """dt = tfactRK * dx / c1 in single precision (kernel/KernelInterface.cpp:125-127)."""
        return _f32(_f32(_f32(self.tfactRK) * _f32(self.gridSpacing)) / _f32(self.c1))

    def GetWindowSize(self):
        """kernel/KernelInterface.cpp:129-134 (C round: half away from zero)."""
        def cround(x):
            return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)
        pe = _f32(self.patcherror)
        tmp = int(cround((pe * 0.7 - 17) / 2.0))
        return int(cround(tmp * 2))

    @staticmethod
    def from_c(s):
        out = PSTDSettings()
        out.calctime, out.c1, out.ampMax, out.freqMax = s.calctime, s.c1, s.ampmax, s.freqmax
        out.rho, out.patcherror, out.tfactRK, out.gridSpacing = s.rho, s.patcherror, s.tfactRK, s.grid_spacing
        out.band_width, out.wave_length = s.band_width, s.wave_length
        out.rk_coefficients = [s.rk_coefficients[i] for i in range(6)]
        out.spectral_interpolation = bool(s.spectral_interpolation)
        out.PMLCells, out.SaveNth = s.pml_cells, s.save_nth
        return out

    def to_c(self):
        s = T.pstd_settings()
        s.calctime = self.calctime
        s.c1 = self.c1
        s.ampmax = self.ampMax
        s.freqmax = self.freqMax
        s.rho = self.rho
        s.patcherror = self.patcherror
        s.tfactRK = self.tfactRK
        s.grid_spacing = self.gridSpacing
        s.band_width = self.band_width
        s.wave_length = self.wave_length
        if len(self.rk_coefficients) != 6:
            raise ValueError("the RK loop uses exactly 6 coefficients")
        for i, c in enumerate(self.rk_coefficients):
            s.rk_coefficients[i] = c
        s.spectral_interpolation = int(self.spectral_interpolation)
        s.pml_cells = self.PMLCells
        s.save_nth = self.SaveNth
        return s


class DomainConfEdge:
    """kernel/KernelInterface.h:163-176."""

    def __init__(self, Absorption=0.0, LR=False):
        self.Absorption = float(Absorption)
        self.LR = bool(LR)


class DomainConf:
    """kernel/KernelInterface.h:181-206; TopLeft/Size are (x, y) in metres."""

    def __init__(self, TopLeft=(0.0, 0.0), Size=(0.0, 0.0)):
        self.TopLeft = tuple(float(v) for v in TopLeft)
        self.Size = tuple(float(v) for v in Size)
        self.T = DomainConfEdge()
        self.L = DomainConfEdge()
        self.B = DomainConfEdge()
        self.R = DomainConfEdge()

    def _edge(self, side):
        return {PSTD_DOMAIN_SIDE_TOP: self.T, PSTD_DOMAIN_SIDE_BOTTOM: self.B,
                PSTD_DOMAIN_SIDE_LEFT: self.L, PSTD_DOMAIN_SIDE_RIGHT: self.R}.get(side)

    def GetAbsorption(self, side):
        e = self._edge(side)
        return e.Absorption if e else 0.0

    def GetLR(self, side):
        e = self._edge(side)
        return e.LR if e else False

    def SetAbsorption(self, sides, absorption):
        for bit in (PSTD_DOMAIN_SIDE_TOP, PSTD_DOMAIN_SIDE_BOTTOM, PSTD_DOMAIN_SIDE_LEFT, PSTD_DOMAIN_SIDE_RIGHT):
            if sides & bit:
                self._edge(bit).Absorption = float(absorption)

    def SetLR(self, sides, lr):
        for bit in (PSTD_DOMAIN_SIDE_TOP, PSTD_DOMAIN_SIDE_BOTTOM, PSTD_DOMAIN_SIDE_LEFT, PSTD_DOMAIN_SIDE_RIGHT):
            if sides & bit:
                self._edge(bit).LR = bool(lr)

    def to_c(self):
        d = T.pstd_domain_conf()
        d.top_left[0], d.top_left[1] = self.TopLeft
        d.size[0], d.size[1] = self.Size
        for name in "TLBR":
            e = getattr(self, name)
            ce = getattr(d, name)
            ce.absorption = e.Absorption
            ce.lr = int(e.LR)
        return d


class PSTDConfiguration:
    """kernel/KernelInterface.h:214-241.

    ``receiver_mode`` selects how receivers read p0: nearest cell
    (abi_types.PSTD_RECEIVER_PHYSICAL or PSTD_RECEIVER_REFERENCE) or spectral
    interpolation at the exact location (PSTD_RECEIVER_SPECTRAL).
    """

    def __init__(self):
        self.Settings = PSTDSettings()
        self.Speakers = []
        self.Receivers = []
        self.Domains = []
        self.receiver_mode = T.PSTD_RECEIVER_PHYSICAL

    @staticmethod
    def CreateDefaultConf():
        """kernel/KernelInterface.cpp:196-249."""
        conf = PSTDConfiguration()
        s = conf.Settings
        s.SetRenderTime(1.0)
        sound_speed = 340.0
        grid_spacing = _f32(0.2)
        s.SetSoundSpeed(sound_speed)
        s.SetAttenuationOfPMLCells(20000.0)
        s.SetDensityOfAir(1.2)
        s.SetPatchError(70.0)
        s.SetFactRK(0.5)
        s.SetGridSpacing(grid_spacing)
        s.SetPMLCells(50)
        s.SetSaveNth(1)
        # (float)(3 * 1e-6 * (c/dx) * (c/dx)): c/dx is a float division, the rest double
        q = _f32(sound_speed / grid_spacing)
        s.SetBandWidth(_f32(3 * 1e-6 * q * q))
        s.SetRKCoefficients(RK_COEFFICIENTS)
        s.SetSpectralInterpolation(True)
        conf.Speakers.append((4.0, 5.0, 0.0))
        conf.Receivers.append((6.0, 5.0, 0.0))
        d = DomainConf((0, 0), (10, 15))
        d.T.Absorption, d.B.Absorption, d.L.Absorption, d.R.Absorption = 1, 0, 0.3, 1
        d2 = DomainConf((10, 4), (20, 15))
        d2.T.Absorption, d2.B.Absorption, d2.L.Absorption, d2.R.Absorption = 1, 0, 1, 1
        conf.Domains += [d, d2]
        return conf

    @staticmethod
    def CreateEmptyConf():
        """kernel/KernelInterface.cpp:251-269."""
        conf = PSTDConfiguration()
        s = conf.Settings
        for setter in (s.SetRenderTime, s.SetSoundSpeed, s.SetAttenuationOfPMLCells, s.SetDensityOfAir,
                       s.SetPatchError, s.SetFactRK, s.SetGridSpacing, s.SetBandWidth):
            setter(0)
        s.SetPMLCells(0)
        s.SetSaveNth(0)
        s.SetSpectralInterpolation(False)
        return conf

    def num_steps(self):
        """Solver::number_of_time_steps = (int)(calctime / dt) in float (kernel/Solver.cpp:41)."""
        return int(_f32(_f32(self.Settings.GetRenderTime()) / self.Settings.GetTimeStep()))

    @staticmethod
    def from_c(c):
        """The configuration a pstd_config describes (arrays copied)."""
        conf = PSTDConfiguration()
        conf.Settings = PSTDSettings.from_c(c.settings)
        conf.Speakers = [tuple(c.speakers[3 * i + j] for j in range(3)) for i in range(c.n_speakers)]
        conf.Receivers = [tuple(c.receivers[3 * i + j] for j in range(3)) for i in range(c.n_receivers)]
        for i in range(c.n_domains):
            cd = c.domains[i]
            d = DomainConf((cd.top_left[0], cd.top_left[1]), (cd.size[0], cd.size[1]))
            for name in "TLBR":
                e, ce = getattr(d, name), getattr(cd, name)
                e.Absorption, e.LR = ce.absorption, bool(ce.lr)
            conf.Domains.append(d)
        conf.receiver_mode = c.receiver_mode
        return conf

    def to_c(self):
        """Returns (pstd_config, keepalive) — keep the second item alive while the first is used."""
        c = T.pstd_config()
        c.settings = self.Settings.to_c()
        spk = (C.c_float * (3 * max(1, len(self.Speakers))))()
        for i, p in enumerate(self.Speakers):
            for j in range(3):
                spk[3 * i + j] = p[j] if j < len(p) else 0.0
        rcv = (C.c_float * (3 * max(1, len(self.Receivers))))()
        for i, p in enumerate(self.Receivers):
            for j in range(3):
                rcv[3 * i + j] = p[j] if j < len(p) else 0.0
        doms = (T.pstd_domain_conf * max(1, len(self.Domains)))()
        for i, d in enumerate(self.Domains):
            doms[i] = d.to_c()
        c.n_speakers = len(self.Speakers)
        c.speakers = C.cast(spk, C.POINTER(C.c_float))
        c.n_receivers = len(self.Receivers)
        c.receivers = C.cast(rcv, C.POINTER(C.c_float))
        c.n_domains = len(self.Domains)
        c.domains = C.cast(doms, C.POINTER(T.pstd_domain_conf))
        c.receiver_mode = int(self.receiver_mode)
        return c, (spk, rcv, doms)
