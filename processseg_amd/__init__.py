"""processseg_amd: processSEG's compiled-pipeline panel hot path on MI355X.

The panel path (batched mixed-radix real FFT, zero-phase FFT filters with the
reference's edge handling, demean, AGC, normalization and vibroseis
correlation, the mutes, time-variant gain curves and calibration) runs in hand-written gfx950 kernels behind the C ABI of
include/pseg.h (libpseg.so).  This package mirrors the reference's Nim API
over that ABI: initPipeline / compile / PipelinePlan.apply and
initCorrelationPlan / correlateInto, plus PipelineBatch (many records with
their own plans as one persistent launch) and Server (zoom windows of a
device-resident survey served by continuous batching).
"""
from ._lib import HipError, UnsupportedOp, WorkspaceError, build, lib  # noqa: F401
from .pipeline import (  # noqa: F401
    ChunkMeta, Corners, FilterSpec, Inf, MuteCurve, NaN, Op, Pipeline, PipelinePlan, agcMeanAbs,
    agcOp, lmoMute, mkLmo, mkTable, mkTime, muteBottomOp, muteTopOp, tableMute, timeMute,
    GainCurve, GainKnot, exponentialGain, gainCurveOp, gcExponential, gcKnots, gcPower,
    heldGain, knotGain, powerGain, ipMinimum, ipZero,
    agcRms, bandPass, bandPassOp, butterBandPass, butterHighCut, butterLowCut, clipOp,
    decodeWindow, demeanOp, detrendOp, filterOp, float32, float64, gainDbOp, highCut, iirOp,
    initPipeline, lowCut, nextFastLen, normalizeOp, normMax, normPercentile, normRms, perPanel,
    perTrace, polarityOp, sameInterval, scaleOp, trLinear, trMean, trNone, tsAuxiliary, tsDead,
    tsLive, tsUnknown, calibrateOp, uncalibrateOp)
from .calibrate import (  # noqa: F401
    Calibration, CalibrationPlan, CalibrationRule, CalibrationSummary, Descale, FileScale,
    RowCalibration, RuleMatch, Sensitivity, StationKey, auAcceleration, auForce, auPressure,
    auRecorded, auVelocity, auVolt, csConverted, csKept, csUnresolved, ctPhysical, ctVolts,
    descale, duMicroVPerCount, duMilliVPerCount, duNanoVPerCount, duVPerCount, ffFallback,
    ffFirst, ffIgnore, fsFile, fsNone, fsRule, initCalibration, initCalibrationPlan,
    initFileScale, openCircuitSensitivity, quantityOf, ratioDbInto, scaleRows, sensitivity,
    stationKey, suDbReVPerMicroPa, suMicroVPerPa, suVPerBar, suVPerG, suVPerInPerS, suVPerKN,
    suVPerLbf, suVPerMPerS, suVPerMPerS2, suVPerN, suVPerPa)
from .correlate import (  # noqa: F401
    CorrelationPlan, Pilot, SweepSpec, SweepTaper, cmCorrelate, cmDeconvolve, generatePilot,
    initCorrelationPlan, initSweepTaper, linearSweep, samplesInLength, sweepSamples)
from .batch import BatchItem, PipelineBatch  # noqa: F401
from .spectrum import SpectrumPlan, initSpectrumPlan  # noqa: F401
from .hilbert import HilbertPlan, initHilbertPlan  # noqa: F401
from .spectrogram import SpectrogramPlan, initSpectrogramPlan  # noqa: F401
from .fk import FkAxes, FkPlan, fkRows, initFkPlan  # noqa: F401
from .fold import (  # noqa: F401
    Binning, FoldMap, FoldOptions, bnCdp, bnGrid, bnInlineCrossline, bnLine, cdpBinning,
    foldDifference, gridBinning, initFoldMap, inlineCrosslineBinning, lineBinning, traceKeyTable)
from .firstbreak import (  # noqa: F401
    FirstBreakParams, FirstBreakPicks, FirstBreakPlan, initFirstBreakPlan, pmCoppens,
    pmEnergyRatio, pmStaLta)
from .decode import (  # noqa: F401
    decodeSamples, fmtIbm32, fmtIeee32, fmtInt8, fmtInt16, fmtInt32, fmtSegd8036, fmtSegd8038,
    fmtSegd8058, fmtSegd8080)
from .serve import ServeResult, Server  # noqa: F401
