"""Autotuner for the engine's runtime kernel config.

Mirrors the reference's tuning/ package (config space, Driver, Manipulator,
Modeling, Tuner): the Driver runs the engine in-process over the C ABI
instead of rebuilding and launching a binary per trial, and the cost model
is a 3-layer MLP trained by this project's HIP kernels instead of XGBoost.
Every pick is re-measured; exhaustive grid search is the fallback and the
judge of the model's ranking.
"""
from .space import INT_MAX, Config, ConfigSpace, SPACE  # noqa: F401
