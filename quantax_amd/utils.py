"""Small host-side helpers: the log-domain psi container and a metric log.

``LogArray`` mirrors ``quantax/utils/big_array.py:152-402`` for real
positive-sign wavefunctions (value = sign * exp(logabs)); ``DataTracer`` is a
list-backed subset of ``quantax/utils/data.py:8-147``.
"""
from __future__ import annotations

import numpy as np
import torch


class LogArray:
    def __init__(self, sign, logabs):
        self.sign = sign
        self.logabs = logabs

    @property
    def shape(self):
        return tuple(self.logabs.shape)

    def value(self):
        return self.sign * torch.exp(self.logabs)

    def __abs__(self):
        return LogArray(torch.ones_like(self.sign), self.logabs)

    def __truediv__(self, other):
        return LogArray(self.sign / other.sign, self.logabs - other.logabs)

    def __getitem__(self, idx):
        return LogArray(self.sign[idx], self.logabs[idx])

    def __len__(self):
        return self.logabs.shape[0]

    def __repr__(self):
        return f"LogArray(sign={self.sign}, logabs={self.logabs})"


class DataTracer:
    def __init__(self):
        self.data = []
        self.time = []

    def append(self, value, time=None):
        self.data.append(float(value))
        self.time.append(float(len(self.time) if time is None else time))

    def __getitem__(self, i):
        return self.data[i]

    def __len__(self):
        return len(self.data)

    def mean(self, n=None):
        d = self.data if n is None else self.data[-n:]
        return float(np.mean(d))

    def save(self, file):
        np.save(file, np.stack([np.asarray(self.time), np.asarray(self.data)]), allow_pickle=False)
