"""Writes config1_train.csv / config1_test.csv: the toy regression problem of
the reference's usage example (test_example.py:18-26, readme.md:47-59).

200 samples, x1 and x2 drawn with Python's random.randint(1, 100) after
random.seed(12345), y = sin(x1) - 78.8*ln(x2) + 4*x1 + 3.31*x2, split 75/25
with sklearn's train_test_split(random_state=12345).  Values are written with
repr(), which round-trips doubles exactly.
"""
import os
from math import log, sin
from random import randint, seed

from sklearn.model_selection import train_test_split

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    random_state = 12345
    samples = 200
    train_size = 0.75
    seed(random_state)
    X = list(zip([randint(1, 100) for _ in range(samples)], [randint(1, 100) for _ in range(samples)]))
    y = [sin(x1) - 78.8 * log(x2) + 4 * x1 + 3.31 * x2 for x1, x2 in X]
    X_train, X_test, y_train, y_test = train_test_split(X, y, train_size=train_size, test_size=1 - train_size,
                                                        random_state=random_state)
    for name, XX, yy in (("config1_train.csv", X_train, y_train), ("config1_test.csv", X_test, y_test)):
        with open(os.path.join(HERE, name), "w") as f:
            f.write("x1,x2,y\n")
            for (a, b), v in zip(XX, yy):
                f.write("%d,%d,%r\n" % (a, b, v))


if __name__ == "__main__":
    main()
