"""Callback dispatch around the training loop (tedeous/callbacks/callback_list.py).

Only the dispatch is provided: user callbacks with ``on_train_begin``,
``on_epoch_begin``, ``on_epoch_end`` and ``on_train_end`` methods (and an
optional ``set_model``) are called at the same points as in the reference.
"""


class Callback:
    def set_model(self, model):
        self.model = model

    def on_train_begin(self, logs=None):
        pass

    def on_epoch_begin(self, logs=None):
        pass

    def on_epoch_end(self, logs=None):
        pass

    def on_train_end(self, logs=None):
        pass


class CallbackList:
    def __init__(self, callbacks=None, model=None):
        self.callbacks = list(callbacks or [])
        self.model = model
        for cb in self.callbacks:
            if hasattr(cb, 'set_model'):
                cb.set_model(model)

    def _call(self, name):
        for cb in self.callbacks:
            fn = getattr(cb, name, None)
            if fn is not None:
                fn()

    def on_train_begin(self):
        self._call('on_train_begin')

    def on_epoch_begin(self):
        self._call('on_epoch_begin')

    def on_epoch_end(self):
        self._call('on_epoch_end')

    def on_train_end(self):
        self._call('on_train_end')
