"""Monitor interface of the engines (Monitor/Monitor.py:264-300): the
engines only call these no-op hooks; live plotting is not implemented."""


class DummyMonitor:
    objectZoom = 1
    probeZoom = 1
    figureUpdateFrequency = 1000000
    verboseLevel = "low"

    def update_encoder(self, *args, **kwargs):
        pass

    def updateBeamWidth(self, *args, **kwargs):
        pass

    def updatePlot(self, object_estimate, probe_estimate):
        pass

    def getOverlap(self, ind1, ind2, probePixelsize):
        pass

    def initializeVisualisation(self):
        pass

    def initializeMonitors(self):
        pass

    def updateObjectProbeErrorMonitor(self, *args, **kwargs):
        pass

    def updateDiffractionDataMonitor(self, *args, **kwargs):
        pass

    def writeEngineName(self, *args, **kwargs):
        pass
