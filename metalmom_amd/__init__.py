"""metalmom_amd: MetalMom's spectral front end on AMD Instinct MI355X (gfx950).

Drop-in for the reference hot path: ``stft``, ``istft``, ``onset_strength``,
``resample``, ``load``, ``get_duration``, ``get_samplerate``,
``feature.melspectrogram``, ``feature.mfcc``, ``feature.power_to_db`` /
``amplitude_to_db`` keep the
signatures of ``metalmom`` (zakkeown/MetalMom ``python/metalmom``) and call
the same C ABI (``mm_*``) implemented by ``libmetalmom.so`` with HIP kernels.
Additive: ``stft_complex``, ``spectral`` (all options), device-resident
batches (``batch.SpectralBatch``) and a request server (``serve``).
"""
__version__ = "0.1.0"

from .core import stft, stft_complex, istft, spectral, resample, load, get_duration, get_samplerate  # noqa: F401
from .feature import (amplitude_to_db, power_to_db, melspectrogram, mfcc, chroma_stft,  # noqa: F401
                      spectral_centroid, spectral_bandwidth, spectral_contrast, spectral_rolloff,
                      spectral_flatness)
from . import feature  # noqa: F401
from .onset import onset_strength  # noqa: F401
