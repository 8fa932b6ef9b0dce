"""sklearn-style RILS-ROLS estimators over the `rils_rols_cpp` extension.

The interface is the reference's (rils_rols/rils_rols.py:16-188): class
names, constructor arguments and defaults, the line printed on construction,
fit / predict / score / model_string / fit_report_string, predict_proba and
the binary-target check of the classifier, with the same messages.  The
behaviour behind it:

* sample_size == 0 picks the sample share before the real fit (reference
  rils_rols.py:51-85): data of at most 10000 rows use everything; larger data
  try a 10% split (train_test_split with the estimator's random_state) under
  1% of the fit-call budget, and keep that share when the sample's R2 and the
  whole set's R2 of the resulting model differ by less than 0.01; the calls
  and seconds spent come off the budget of the real fit.
* after a regression fit the model string is simplified by sympy under a
  2-second deadline, falling back to sympify (rils_rols.py:116-125).

Two keyword arguments are added at the end of each constructor: `precision`
("fp64", the default and bit-for-bit path, or "fp32", sweeps scored by the
certified FP32 form) and `device` (HIP device index).
"""
import concurrent.futures
import time
import warnings

import numpy as np
import pandas as pd
import sympy
from sklearn.base import BaseEstimator
from sklearn.metrics import accuracy_score, r2_score
from sklearn.model_selection import train_test_split

import rils_rols_cpp

from .utils import binarize, complexity_sympy, proba

warnings.filterwarnings("ignore")

SIMPLIFY_SECONDS = 2.0
AUTO_SAMPLE_MAX_ROWS = 10000        # up to this many rows sample_size=0 means the full data
AUTO_SAMPLE_SHARES = (0.1,)         # shares tried, in order
AUTO_SAMPLE_BUDGET_SHARE = 0.01     # share of max_fit_calls one trial may spend
AUTO_SAMPLE_R2_GAP = 0.01           # a share is representative below this R2 gap


def _matrix(X):
    """Rows x features float64 array from a DataFrame, nested lists or an array."""
    if isinstance(X, pd.DataFrame):
        X = X.to_numpy()
    return np.asarray(X, dtype=np.float64)


def _vector(y):
    if isinstance(y, (pd.DataFrame, pd.Series)):
        y = y.to_numpy()
    return np.asarray(y, dtype=np.float64).reshape(-1)


def _simplify_within(expr, seconds, verbose):
    """sympy.simplify(expr, ratio=1) if it returns within `seconds`, else
    sympify(expr).  The worker is not waited for: a simplification that runs
    long finishes in the background and is discarded."""
    worker = concurrent.futures.ThreadPoolExecutor(max_workers=1)
    try:
        return worker.submit(sympy.simplify, expr, ratio=1).result(timeout=seconds)
    except Exception:
        if verbose:
            print("Simplification failed within given timeout, so just doing sympify.")
        return sympy.sympify(expr)
    finally:
        worker.shutdown(wait=False)


class RILSROLSBase(BaseEstimator):

    def __init__(self, classification=None, max_fit_calls=100000, max_time=100, complexity_penalty=0.001,
                 max_complexity=50, sample_size=1, verbose=False, random_state=0, precision="fp64", device=0):
        print(f'Calling with max_fit_calls={max_fit_calls} max_time={max_time} complexity_penalty={complexity_penalty} '
              f'max_complexity={max_complexity} sample_size={sample_size} verbose={verbose} random_state={random_state}')
        self.classification = classification
        self.max_fit_calls = max_fit_calls
        self.max_time = max_time
        self.complexity_penalty = complexity_penalty
        self.max_complexity = max_complexity
        self.sample_size = sample_size
        self.verbose = verbose
        self.random_state = random_state
        self.precision = precision
        self.device = device
        self.rr_cpp = None
        self.model = None
        self.model_simp = None

    # ------------------------------------------------------------------ fit
    def fit(self, X, y):
        if self.sample_size == 0:
            self._choose_sample_size(X, y)
        elif not 0 < self.sample_size <= 1:
            raise Exception('Sample size parameter must belong to interval (0, 1], while value 0 means it is '
                            'automatically tuned.')
        self.fit_inner(X, y)

    def _choose_sample_size(self, X, y):
        """sample_size == 0: pick the share, charge the trials to the budget."""
        print('Automatically tuning sample size:')
        if len(X) <= AUTO_SAMPLE_MAX_ROWS:
            print('Training size is smaller than 10000 so setting it to full sample 1 (100%).')
            self.sample_size = 1
            return
        started = time.time()
        budget = self.max_fit_calls
        trial_budget = budget * AUTO_SAMPLE_BUDGET_SHARE
        spent = 0
        chosen = 1
        for share in AUTO_SAMPLE_SHARES:
            X_part, _, y_part, _ = train_test_split(X, y, train_size=share, random_state=self.random_state)
            # the trial fits the drawn rows in full under its slice of the budget
            self.max_fit_calls, self.sample_size = trial_budget, 1
            self.fit_inner(X_part, y_part)
            spent += trial_budget
            if self._share_is_representative(share, X_part, y_part, X, y):
                chosen = share
                break
        print(f'Setting sample_size={chosen}')
        self.sample_size = chosen
        self.max_time -= time.time() - started
        self.max_fit_calls = budget - spent

    def _share_is_representative(self, share, X_part, y_part, X, y):
        on_part, on_all = self.predict(X_part), self.predict(X)
        try:
            r2_part, r2_all = r2_score(y_part, on_part), r2_score(y, on_all)
        except Exception:
            print('Error while calculating R2.')
            return False
        print(f'Sample size {share} --> R2={r2_all} R2_sample={r2_part}')
        return abs(r2_all - r2_part) < AUTO_SAMPLE_R2_GAP

    def fit_inner(self, X, y):
        if self.verbose and isinstance(X, pd.DataFrame):
            print('Converting X dataframe to just list of lists.')
        if self.verbose and isinstance(y, pd.DataFrame):
            print('Converting y dataframe to list of lists.')
        Xm, yv = _matrix(X), _vector(y)
        rows, features = Xm.shape
        self.rr_cpp = rils_rols_cpp.rils_rols(self.classification, int(self.max_fit_calls), int(self.max_time),
                                              self.complexity_penalty, self.max_complexity, self.sample_size,
                                              self.verbose, int(self.random_state))
        self.rr_cpp.set_precision(self.precision)
        self.rr_cpp.set_device(int(self.device))
        self.rr_cpp.fit(Xm.reshape(-1), yv, rows, features)
        self.model = self.rr_cpp.get_model_string()
        self.best_time = self.rr_cpp.get_best_time()
        self.total_time = self.rr_cpp.get_total_time()
        self.fit_calls = self.rr_cpp.get_fit_calls()
        if self.classification is True:
            if self.verbose:
                print('Finished optimization, not doing symplification for the classification scenario.')
            self.model_simp = self.model
        else:
            if self.verbose:
                print('Finished otimization, doing final symplification...')
            self.model_simp = _simplify_within(self.model, SIMPLIFY_SECONDS, self.verbose)
        return (self.model, self.model_simp)

    # -------------------------------------------------------------- queries
    def check_model(self):
        if self.model is None or self.rr_cpp is None:
            raise Exception("Cannot predict because model is not build yet. First call fit().")

    def predict(self, X):
        if self.verbose and isinstance(X, pd.DataFrame):
            print('Converting X dataframe to just list of lists.')
        self.check_model()
        Xm = _matrix(X)
        return self.rr_cpp.predict(Xm.reshape(-1), Xm.shape[0], Xm.shape[1])

    def model_string(self):
        self.check_model()
        return self.model_simp

    def fit_report_string(self):
        self.check_model()
        fields = [("maxTime", self.max_time), ("maxFitCalls", self.max_fit_calls), ("seed", self.random_state),
                  ("sizePenalty", self.complexity_penalty), ("maxComplexity", self.max_complexity),
                  ("sampleShare", self.sample_size), ("totalTime", "%.1f" % self.total_time),
                  ("bestTime", self.best_time), ("fitCalls", self.fit_calls),
                  ("simpSize", complexity_sympy(self.model_simp)), ("expr", self.model), ("exprSimp", self.model_simp)]
        return "\t".join("%s=%s" % (k, v) for k, v in fields)


class RILSROLSRegressor(RILSROLSBase):

    def __init__(self, max_fit_calls=100000, max_time=100, complexity_penalty=0.001, max_complexity=50, sample_size=1,
                 verbose=False, random_state=0, precision="fp64", device=0):
        super().__init__(False, max_fit_calls, max_time, complexity_penalty, max_complexity, sample_size, verbose,
                         random_state, precision, device)

    def score(self, X, y):
        return r2_score(y, self.predict(X))


class RILSROLSBinaryClassifier(RILSROLSBase):

    def __init__(self, max_fit_calls=100000, max_time=100, complexity_penalty=0.001, max_complexity=50, sample_size=1,
                 verbose=False, random_state=0, precision="fp64", device=0):
        super().__init__(True, max_fit_calls, max_time, complexity_penalty, max_complexity, sample_size, verbose,
                         random_state, precision, device)

    def check_binary_targets(self, y):
        labels = _vector(y)
        if np.any((labels != 0) & (labels != 1)):
            raise Exception('The classifier works only for binary targets, so allowed target values are 0 or 1.')

    def predict(self, X):
        return binarize(super().predict(X))

    def predict_proba(self, X):
        return proba(self.predict(X))

    def score(self, X, y):
        self.check_binary_targets(y)
        return accuracy_score(y, self.predict(X))

    def fit(self, X, y):
        self.check_binary_targets(y)
        return super().fit(X, y)

