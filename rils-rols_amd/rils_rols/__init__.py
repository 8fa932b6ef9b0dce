"""RILS-ROLS estimators whose candidate-fitness hot path runs on MI355X.

Drop-in for the reference package `rils_rols` (rils_rols/rils_rols.py,
rils_rols/utils.py): same classes, constructor arguments and methods.
"""
from .rils_rols import RILSROLSBase, RILSROLSBinaryClassifier, RILSROLSRegressor  # noqa: F401
