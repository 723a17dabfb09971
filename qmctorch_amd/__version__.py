__version__ = "0.4.0+mi355x.1"
